"""CPU oracle for the learned surrogate (test infrastructure only).

The surrogate has no counterpart in the reference.  It replaces the
acceleration part of the SDPD pair loop (the F sum of
spatialpy/solvers/c_base/ssa_sdpd-c-simulation-engine/src/pairwise_interaction.cpp,
restated in spatialpy_amd/csrc/sdpd.hip k_forces).  That sum is, per pair,

    a_ij dx + b_ij dv + c_ij v_i + d_ij v_j

with scalar coefficients of rotation-invariant quantities (pressure,
viscous and transport-velocity terms).  The network keeps that structure
and learns the coefficients:

    features f_ij (16 invariants, padded to 32)   -> 64 -> 64 -> 4, plus a
                  linear map of the 16 features    (o_ij)
    message   m_ij = u(R) (o0 dx/h + o1 dv/v_s + o2 v_i/v_s + o3 v_j/v_s)
    pred_i    = sum_j m_ij                          (units of acc_scale)

u(R) = (1-R)^2 R / (R + 0.001), R = r/h, is a fixed smooth cut-off envelope
(zero at R = 1).  Hidden activation: hard SiLU hs(z) = z min(max(z+3,0),6)/6.
The prediction is equivariant under rotations and reflections of the frame
by construction.

numpy FP64.  With ``emulate=True`` the two hidden layers' products use
bf16-rounded weights (round to nearest even) and activations split into a
bf16 high part and a bf16 low part (a = hi + lo + O(2^-17 a)), as the GPU's
matrix-core products do with FP32 accumulation; the head (64 -> 4), biases,
envelope and messages stay unrounded (FP32 on the GPU).  The backward pass is pinned by central finite differences of the
FP64 loss (tests/test_oracle_cpu.py).

Parameter layout (flat, FP32 on the GPU): W1 [32][64], b1 [64], W2 [64][64],
b2 [64], W3 [64][4], b3 [4], L [16][4] (the linear skip).
"""
import numpy as np

NF, KIN, HID, NOUT = 16, 32, 64, 4
_SHAPES = [("W1", (KIN, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)), ("W3", (HID, NOUT)),
           ("b3", (NOUT,)), ("L", (NF, NOUT))]
P = sum(int(np.prod(s)) for _, s in _SHAPES)
# scales: rho0, v_scale, rho_scale, nu_scale, acc_scale, tv_scale
NSCALES = 6


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    assert w.size == P, (w.size, P)
    out, o = {}, 0
    for k, s in _SHAPES:
        m = int(np.prod(s))
        out[k] = w[o:o + m].reshape(s)
        o += m
    return out


def pack(p):
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k, _ in _SHAPES])


def bf16(a):
    """Round to bf16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def split(a):
    """bf16 high part plus bf16 low part of FP32(a)."""
    a32 = np.asarray(a, dtype=np.float32).astype(np.float64)
    hi = bf16(a32)
    lo = bf16((a32 - hi).astype(np.float32))
    return hi + lo


def hs(z):
    return z * np.clip(z + 3.0, 0.0, 6.0) / 6.0


def hs_grad(z):
    return np.where(z < -3.0, 0.0, np.where(z > 3.0, 1.0, (2.0 * z + 3.0) / 6.0))


def envelope(R):
    q = 1.0 - R
    return q * q * R / (R + 0.001)


def edges(g, sel):
    """Edges (slot order) of the selected particles: (p, i, j, rd) with p the
    position of i in `sel`.  Every slot k < deg[i] is an edge; those with
    rd/h > 1 or rd == 0 carry no message (the pair loop skips them)."""
    deg, nidx, nr = g["deg"], g["nidx"], g["nr"]
    P_, I, J, RD = [], [], [], []
    for p, i in enumerate(sel):
        for k in range(int(deg[i])):
            P_.append(p)
            I.append(i)
            J.append(int(nidx[k, i]))
            RD.append(float(nr[k, i]))
    return np.array(P_, np.int64), np.array(I, np.int64), np.array(J, np.int64), np.array(RD)


def features(g, sc, I, J, RD):
    """(X [E][32], basis [E][4][3], u [E]) of the edges, rounded to FP32 as
    the GPU holds them.  With r = rho/rho0 and n_j = m_j/m_i, features:
      0 R = r/h   1 (1/(r_i r_j) - 1)/rho_s (the viscous term's density
        factor, model.cpp:118)   2, 3 (r - 1)/rho_s of i, j
      4 sigma: the pair loop's pressure sign rule is taken
        (P_i/rho_i^2 + P_j/rho_j^2 < 0, model.cpp:111-112)
      5 dv.dx/(h v_s)   6, 7 (v - vt).dx/(h tv_s) of i, j
      8 (s P_i/rho_i^2 + P_j/rho_j^2) in units of P0 rho_s/rho0^2, s = -1
        when sigma, else 1
      9 n_j - 1   10, 11 nu/nu_s of i, j   12, 13 solid flags of i, j
      14, 15 features 6, 7 times r_i resp. r_j times
        V = (1/r_i^2 + n_j^2/r_j^2)/2 (the transport tensor's factors,
        model.cpp:127-130)"""
    rho0, vs, rs, ns, _, ts = [float(s) for s in sc]
    h = float(g["h"])
    x, v, vt = g["x"], g["v"], g["vt"]
    rho, mass, nu, solid = g["rho"], g["mass"], g["nu"], g["solid"]
    E = len(I)
    X = np.zeros((E, KIN))
    B = np.zeros((E, 4, 3))
    u = np.zeros(E)
    for e in range(E):
        i, j, rd = I[e], J[e], RD[e]
        if rd / h > 1.0 or rd == 0.0:
            continue
        R = rd / h
        d = (x[i] - x[j]) / h
        dv = (v[i] - v[j]) / vs
        vi, vj = v[i] / vs, v[j] / vs
        ri, rj = rho[i] / rho0, rho[j] / rho0
        sigma = (ri - 1.0) * rj * rj + (rj - 1.0) * ri * ri < 0.0
        pgi, pgj = (ri - 1.0) / ri ** 2, (rj - 1.0) / rj ** 2
        nj = mass[j] / mass[i]
        V = 0.5 * (1.0 / ri ** 2 + nj ** 2 / rj ** 2)
        ti, tj = ((v[i] - vt[i]) @ d) / ts, ((v[j] - vt[j]) @ d) / ts
        f = [R, (1.0 / (ri * rj) - 1.0) / rs, (ri - 1.0) / rs, (rj - 1.0) / rs, float(sigma), dv @ d, ti, tj,
             ((-pgi if sigma else pgi) + pgj) / rs, nj - 1.0, nu[i] / ns, nu[j] / ns, float(solid[i] != 0),
             float(solid[j] != 0), ti * ri * V, tj * rj * V]
        X[e, :NF] = np.float32(f)
        B[e] = np.float32([d, dv, vi, vj])
        u[e] = np.float32(envelope(np.float32(R)))
    return X, B, u


def forward(w, g, sc, sel, emulate=True):
    """pred [len(sel)][3] in units of acc_scale, and the cache."""
    p = unpack(w)
    mmw = (lambda W: bf16(W)) if emulate else (lambda W: W)
    act = split if emulate else (lambda a: a)
    Pi, I, J, RD = edges(g, sel)
    X, B, u = features(g, sc, I, J, RD)
    Z1 = act(X) @ mmw(p["W1"]) + p["b1"]
    H1 = hs(Z1)
    Z2 = act(H1) @ mmw(p["W2"]) + p["b2"]
    H2 = hs(Z2)
    # the head and the linear skip from the features stay FP32 on the GPU
    O = H2 @ p["W3"] + X[:, :NF] @ p["L"] + p["b3"]
    M = u[:, None] * np.einsum("ek,ekc->ec", O, B)
    pred = np.zeros((len(sel), 3))
    np.add.at(pred, Pi, M)
    c = dict(P=Pi, X=X, B=B, u=u, Z1=Z1, H1=H1, Z2=Z2, H2=H2, O=O, act=act, mmw=mmw)
    return pred, c


def loss_grad(w, g, sc, sel, y, emulate=False):
    """L = (1/n_sel) sum_i |pred_i - y_i|^2 (y in acc_scale units) and dL/dw."""
    p = unpack(w)
    pred, c = forward(w, g, sc, sel, emulate)
    act, mmw = c["act"], c["mmw"]
    r = pred - np.asarray(y, dtype=np.float64)
    ns = len(sel)
    loss = float((r * r).sum() / ns)
    dpred = 2.0 * r / ns
    dM = dpred[c["P"]]
    dO = c["u"][:, None] * np.einsum("ekc,ec->ek", c["B"], dM)
    gr = {"W3": c["H2"].T @ dO, "b3": dO.sum(0), "L": c["X"][:, :NF].T @ dO}
    dH2 = dO @ p["W3"].T
    dZ2 = dH2 * hs_grad(c["Z2"])
    gr["W2"] = act(c["H1"]).T @ dZ2
    gr["b2"] = dZ2.sum(0)
    dH1 = dZ2 @ mmw(p["W2"]).T
    dZ1 = dH1 * hs_grad(c["Z1"])
    gr["W1"] = act(c["X"]).T @ dZ1
    gr["b1"] = dZ1.sum(0)
    return loss, pack(gr)


def init_params(seed=0):
    """He-scaled random parameters (for tests): W1 rows past the 16 features
    are zero, W3 small."""
    rng = np.random.default_rng(seed)
    p = {"W1": np.zeros((KIN, HID)), "b1": rng.normal(0, 0.1, HID),
         "W2": rng.normal(0, np.sqrt(2.0 / HID), (HID, HID)), "b2": rng.normal(0, 0.1, HID),
         "W3": rng.normal(0, 0.3 / np.sqrt(HID), (HID, NOUT)), "b3": rng.normal(0, 0.05, NOUT),
         "L": rng.normal(0, 0.2 / np.sqrt(NF), (NF, NOUT))}
    p["W1"][:NF] = rng.normal(0, np.sqrt(2.0 / NF), (NF, HID))
    return pack(p)


def random_graph(n=200, seed=0, h=0.12, dim=3, solid_frac=0.1):
    """A particle state in the unit box with ELL lists of every pair within
    h (slot order: by index), densities near rho0 = 1, random velocities,
    transport velocities, masses, viscosities and a few solid particles."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3))
    x[:, :dim] = rng.uniform(0, 1, (n, dim))
    v = np.zeros((n, 3))
    v[:, :dim] = rng.normal(0, 0.5, (n, dim))
    vt = v.copy()
    vt[:, :dim] += rng.normal(0, 0.05, (n, dim))
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    nb = [np.nonzero((d[i] <= h) & (np.arange(n) != i))[0] for i in range(n)]
    W = max(1, max(len(q) for q in nb))
    deg = np.array([len(q) for q in nb], np.int32)
    nidx = np.zeros((W, n), np.int32)
    nr = np.zeros((W, n))
    for i, q in enumerate(nb):
        nidx[:len(q), i] = q
        nr[:len(q), i] = d[i, q]
    return {"n": n, "W": W, "dim": dim, "h": h, "x": x, "v": v, "vt": vt,
            "rho": 1.0 + rng.normal(0, 0.02, n), "mass": rng.uniform(0.8, 1.2, n) * 1e-3,
            "nu": rng.uniform(0.5, 1.5, n) * 1e-2, "solid": (rng.uniform(0, 1, n) < solid_frac).astype(np.int32),
            "deg": deg, "nidx": nidx, "nr": nr}


def default_scales(g):
    v = g["v"]
    return np.array([1.0, np.sqrt((v * v).sum(1).mean()), 0.02, 1e-2, 1.0,
                     max(np.sqrt(((g["v"] - g["vt"]) ** 2).sum(1).mean()), 1e-6)])


def sdpd_pair_terms(x, v, vt, rho, mass, nu, rp, idx, dist, dwdr, h, P0, rho0, dim):
    """The pair loop's acceleration terms as coefficients of the four basis
    vectors: for every listed pair (i, j) with 0 < r <= h, (a, b, c, d) with
    pair term a dx + b dv + c v_i + d v_j (model.cpp:107-131: pressure force
    fp with its sign rule, viscous force fv, transport-tensor force ft, whose
    tensor 0.5 (rho_i v_i (vt_i - v_i)^T + rho_j v_j (vt_j - v_j)^T) applied
    to dx is c v_i + d v_j).  Returns (i, j, coef [E][4])."""
    I, J, CO = [], [], []
    n = len(rho)
    for i in range(n):
        Pi = P0 * (rho[i] / rho0 - 1.0)
        for e in range(int(rp[i]), int(rp[i + 1])):
            j, r, dW = int(idx[e]), float(dist[e]), float(dwdr[e])
            if r / h > 1.0 or r == 0.0:
                continue
            dx = np.zeros(3)
            dx[:dim] = x[i, :dim] - x[j, :dim]
            Pj = P0 * (rho[j] / rho0 - 1.0)
            pg = Pi / rho[i] ** 2 + Pj / rho[j] ** 2
            if pg < 0:
                pg = -Pi / rho[i] ** 2 + Pj / rho[j] ** 2
            wf = dW / (r + 0.001 * h)
            a = -mass[j] * pg * wf
            b = mass[j] * (2.0 * nu[i] * nu[j] / (nu[i] + nu[j])) * wf / (rho[i] * rho[j])
            vsum = (mass[i] / rho[i]) ** 2 + (mass[j] / rho[j]) ** 2
            k = vsum * wf / mass[i] * 0.5
            c = k * rho[i] * ((vt[i] - v[i]) @ dx)
            d = k * rho[j] * ((vt[j] - v[j]) @ dx)
            I.append(i)
            J.append(j)
            CO.append((a, b, c, d))
    return np.array(I, np.int64), np.array(J, np.int64), np.array(CO).reshape(-1, 4)
