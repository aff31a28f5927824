"""CPU oracle for the message-passing surrogate (test infrastructure only).

The network has no counterpart in the reference; this file restates it in
numpy FP64: forward, mean-squared-error loss and its gradient with respect
to every parameter.  With ``emulate_bf16=True`` every matrix product rounds
its two operands to bf16 (round to nearest even) and accumulates in FP64,
as the GPU's matrix-core products do with FP32 accumulation; everything
else (edge sums, SiLU, decoder, bias sums) stays FP64.  The backward pass is
pinned by finite differences of the FP64 loss (tests/test_oracle_cpu.py).

Parameter layout (flat, FP32 on the GPU): We [32][64], be [64]; per layer
W1ab [64][128], W1c [4][64], b1 [64], W2 [64][64], b2 [64], Wn1 [128][64],
bn1 [64], Wn2 [64][64], bn2 [64]; Wd [64][3], bd [3].
"""
import numpy as np

HID, LAYERS, NIN, NOUT = 64, 3, 32, 3
_L = [("W1ab", (HID, 128)), ("W1c", (4, HID)), ("b1", (HID,)), ("W2", (HID, HID)), ("b2", (HID,)),
      ("Wn1", (128, HID)), ("bn1", (HID,)), ("Wn2", (HID, HID)), ("bn2", (HID,))]
LSIZE = sum(int(np.prod(s)) for _, s in _L)
P = NIN * HID + HID + LAYERS * LSIZE + HID * NOUT + NOUT


def unpack(w):
    w = np.asarray(w, dtype=np.float64)
    assert w.size == P
    out, o = {}, 0

    def take(shape):
        nonlocal o
        m = int(np.prod(shape))
        v = w[o:o + m].reshape(shape)
        o += m
        return v
    out["We"], out["be"] = take((NIN, HID)), take((HID,))
    out["layers"] = [{k: take(s) for k, s in _L} for _ in range(LAYERS)]
    out["Wd"], out["bd"] = take((HID, NOUT)), take((NOUT,))
    return out


def pack(p):
    parts = [p["We"].ravel(), p["be"].ravel()]
    for lay in p["layers"]:
        parts += [lay[k].ravel() for k, _ in _L]
    parts += [p["Wd"].ravel(), p["bd"].ravel()]
    return np.concatenate(parts)


def bf16(a):
    """Round to bf16 (nearest even) and return as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return u.view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s + z * s * (1.0 - s)


def _edges(deg, nidx):
    n = deg.size
    src, dst = [], []
    for k in range(nidx.shape[0]):
        sel = np.nonzero(deg > k)[0]
        src.append(sel)
        dst.append(nidx[k, sel])
    return np.concatenate(src), np.concatenate(dst)


def _edge_feat(x, inv_h, i, j):
    x = np.asarray(x, dtype=np.float32).astype(np.float64)
    d = (x[i] - x[j]) * np.float64(np.float32(inv_h))
    r = np.sqrt((d * d).sum(1))
    return np.column_stack([d, r])


def forward(w, x, X0, deg, nidx, inv_h, emulate_bf16=False):
    """Returns (out [n][3], cache)."""
    p = unpack(w)
    mm = (lambda a, b: bf16(a) @ bf16(b)) if emulate_bf16 else (lambda a, b: a @ b)
    X0 = np.asarray(X0, dtype=np.float64)
    deg = np.asarray(deg)
    i, j = _edges(deg, np.asarray(nidx))
    e = _edge_feat(x, inv_h, i, j)
    n = X0.shape[0]
    c = {"i": i, "j": j, "e": e, "X0": X0, "deg": deg.astype(np.float64), "mm": mm}
    c["Ze"] = mm(X0, p["We"]) + p["be"]
    H = silu(c["Ze"])
    c["H"] = [H]
    c["lay"] = []
    for lay in p["layers"]:
        AB = mm(H, lay["W1ab"])
        z = AB[i, :HID] + AB[j, HID:] + e @ lay["W1c"] + lay["b1"]
        S = np.zeros((n, HID))
        np.add.at(S, i, silu(z))
        M = mm(S, lay["W2"]) + c["deg"][:, None] * lay["b2"]
        HM = np.concatenate([H, M], axis=1)
        Z1 = mm(HM, lay["Wn1"]) + lay["bn1"]
        U = silu(Z1)
        Hn = H + mm(U, lay["Wn2"]) + lay["bn2"]
        c["lay"].append({"AB": AB, "z": z, "S": S, "M": M, "HM": HM, "Z1": Z1, "U": U})
        H = Hn
        c["H"].append(H)
    out = H @ p["Wd"] + p["bd"]
    return out, c


def loss_grad(w, x, X0, deg, nidx, inv_h, y, emulate_bf16=False):
    """Mean squared error over n*3 outputs and its gradient (flat, P)."""
    p = unpack(w)
    out, c = forward(w, x, X0, deg, nidx, inv_h, emulate_bf16)
    mm = c["mm"]
    n = out.shape[0]
    r = out - np.asarray(y, dtype=np.float64)
    loss = float((r * r).sum() / (3 * n))
    dout = 2.0 * r / (3 * n)
    g = {"layers": [dict() for _ in range(LAYERS)]}
    H3 = c["H"][-1]
    g["Wd"] = H3.T @ dout
    g["bd"] = dout.sum(0)
    dH = dout @ p["Wd"].T
    i, j, e = c["i"], c["j"], c["e"]
    for l in range(LAYERS - 1, -1, -1):
        lay, ca, gl = p["layers"][l], c["lay"][l], g["layers"][l]
        H = c["H"][l]
        dU = mm(dH, lay["Wn2"].T)
        gl["Wn2"] = mm(ca["U"].T, dH)
        gl["bn2"] = dH.sum(0)
        dZ1 = dU * silu_grad(ca["Z1"])
        dHM = mm(dZ1, lay["Wn1"].T)
        gl["Wn1"] = mm(ca["HM"].T, dZ1)
        gl["bn1"] = dZ1.sum(0)
        dM = dHM[:, HID:]
        dS = mm(dM, lay["W2"].T)
        gl["W2"] = mm(ca["S"].T, dM)
        gl["b2"] = (c["deg"][:, None] * dM).sum(0)
        gz = dS[i] * silu_grad(ca["z"])
        dAB = np.zeros((n, 128))
        np.add.at(dAB[:, :HID], i, gz)
        np.add.at(dAB[:, HID:], j, gz)
        gl["W1c"] = e.T @ gz
        gl["b1"] = gz.sum(0)
        gl["W1ab"] = mm(H.T, dAB)
        dH = dH + dHM[:, :HID] + mm(dAB, lay["W1ab"].T)
    dZe = dH * silu_grad(c["Ze"])
    g["We"] = mm(c["X0"].T, dZe)
    g["be"] = dZe.sum(0)
    return loss, pack(g)


def init_params(seed=0):
    """A Glorot-scaled random parameter vector (for tests)."""
    rng = np.random.default_rng(seed)
    p = {"We": np.zeros((NIN, HID)), "be": rng.normal(0, 0.1, HID), "layers": []}
    p["We"][:10] = rng.uniform(-1, 1, (10, HID)) * np.sqrt(6 / (10 + HID))
    for _ in range(LAYERS):
        a = np.sqrt(6 / (2 * HID + 4 + HID))
        p["layers"].append({"W1ab": rng.uniform(-a, a, (HID, 128)), "W1c": rng.uniform(-a, a, (4, HID)),
                            "b1": rng.normal(0, 0.1, HID),
                            "W2": rng.uniform(-1, 1, (HID, HID)) * np.sqrt(6 / (2 * HID)),
                            "b2": rng.normal(0, 0.1, HID),
                            "Wn1": rng.uniform(-1, 1, (128, HID)) * np.sqrt(6 / (3 * HID)),
                            "bn1": rng.normal(0, 0.1, HID),
                            "Wn2": rng.uniform(-1, 1, (HID, HID)) * 0.5 * np.sqrt(6 / (2 * HID)),
                            "bn2": rng.normal(0, 0.1, HID)})
    p["Wd"] = rng.uniform(-1, 1, (HID, NOUT)) * 0.3
    p["bd"] = rng.normal(0, 0.1, NOUT)
    return pack(p)


def random_graph(n=200, seed=0, h=0.25):
    """Points in the unit square with ELL lists of all pairs within h."""
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 3), np.float32)
    x[:, :2] = rng.uniform(0, 1, (n, 2))
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    nb = [np.nonzero((d[i] <= h) & (np.arange(n) != i))[0] for i in range(n)]
    W = max(1, max(len(v) for v in nb))
    deg = np.array([len(v) for v in nb], np.int32)
    nidx = np.zeros((W, n), np.int32)
    for i, v in enumerate(nb):
        nidx[:len(v), i] = v
    X0 = np.zeros((n, NIN), np.float32)
    X0[:, :10] = rng.normal(0, 1, (n, 10))
    X0[:, 6] = 1.0
    y = rng.normal(0, 1, (n, NOUT)).astype(np.float32)
    return x, X0, deg, nidx, np.float32(1.0 / h), y
