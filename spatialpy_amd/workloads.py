"""Synthetic workloads at the benchmark configurations (and scaled-down
variants for tests), lowered straight to SystemArrays.

Domain building follows the reference's semantics: CartesianLattice fill
order x-outer, y, z-inner with per-axis rounding (spatialpy/core/lattice.py:
167-231), ``create_2D_domain``/``create_3D_domain`` volumes and densities
(core/domain.py:508-645), ``find_h`` = 2.2 x the largest nearest-neighbour
distance (core/domain.py:751-769), dimension detection (solvers/solver.py:
407-413), TimeSpan output steps (core/timespan.py:72-85).
"""
import math

import numpy as np

from .lowering import SystemArrays, K_ZERO, K_FIRST, K_AB, K_2X


def _digits(*vals):
    out = 0
    for v in vals:
        s = str(v)
        if "." in s:
            out = max(out, len(s.split(".")[1]))
    return out


def cartesian_lattice(xmin, xmax, dx, ymin, ymax, dy, zmin=0.0, zmax=0.0, dz=0.0, inside=None):
    """Points of CartesianLattice.apply (core/lattice.py:167-219)."""
    # Lattice.__init__ adds a float center [0.0]*3 (core/lattice.py:27-35)
    xmin, xmax, ymin, ymax = xmin + 0.0, xmax + 0.0, ymin + 0.0, ymax + 0.0
    zmin, zmax = zmin + 0.0, zmax + 0.0
    xd = _digits(dx, xmin, xmax)
    yd = _digits(dy, ymin, ymax)
    zd = _digits(dz, zmin, zmax)
    pts = []
    for x in np.arange(xmin, xmax + dx, dx):
        x = round(x, xd)
        for y in np.arange(ymin, ymax + dy, dy):
            y = round(y, yd)
            if dz == 0:
                zs = [0.0]
            else:
                zs = [round(z, zd) for z in np.arange(zmin, zmax + dz, dz)]
            for z in zs:
                if inside is None or inside(x, y, z):
                    pts.append((float(x), float(y), float(z)))
    return np.array(pts, dtype=np.float64).reshape(-1, 3)


def find_h(x):
    """core/domain.py:751-769 (scipy KDTree, 2 nearest incl. self)."""
    from scipy.spatial import KDTree
    d, _ = KDTree(x).query(x, 2)
    return 2.2 * max(d[:, 1])


def detect_dimension(x):
    """solvers/solver.py:407-413."""
    if not np.count_nonzero(x[:, 1]):
        return 1
    if not np.count_nonzero(x[:, 2]):
        return 2
    return 3


def timespan_steps(t, num_points, timestep_size):
    """TimeSpan.linspace + _set_timesteps (core/timespan.py:72-85, 90-130)."""
    items = np.linspace(0, t, num_points)
    output_interval = items[1] - items[0]
    num_steps = len(items) - 1
    output_freq = output_interval / timestep_size
    num_timesteps = math.ceil(num_steps * output_freq)
    steps = np.arange(0, num_timesteps + timestep_size, output_freq)
    steps = np.unique(np.round(steps).astype(int))
    return num_timesteps, steps.astype(np.uint32)


def scatter(rng, n_vox, count, allowed=None):
    """ScatterInitialCondition.apply (core/initialcondition.py:160-200),
    with an explicit generator instead of the global numpy.random."""
    out = np.zeros(n_vox, dtype=np.int64)
    if allowed is None:
        idx = rng.integers(0, n_vox, size=count)
    else:
        allowed = np.asarray(allowed)
        idx = allowed[rng.integers(0, len(allowed), size=count)]
    np.add.at(out, idx, 1)
    return out


def bounding_box(x):
    """Domain limits after a fill (Lattice._update_limits, core/lattice.py:37-44)."""
    lo, hi = x.min(axis=0), x.max(axis=0)
    return (lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])


def _sys_from_lattice(x, vol, mass=1.0, nu=1.0, type_id=1, fixed=True):
    n = x.shape[0]
    return dict(x=x, type=np.full(n, type_id, np.int32), mass=np.full(n, mass),
                rho=np.full(n, mass / vol), nu=np.full(n, nu),
                fixed=np.full(n, int(fixed), np.int32))


def birth_death_2d(numx=10, numy=10, count=100, D=0.1, k_birth=10.0, k_death=0.1,
                   t_end=10, num_points=11, dt=1.0, seed=0):
    """Config 1: the reference's system-test model test/models/birth_death.py:19-50
    (10x10 fixed particles on [0,1]^2, D=0.1, 100 scattered, birth/death)."""
    xl = np.linspace(0, 1, numx)
    yl = np.linspace(0, 1, numy)
    x = cartesian_lattice(0, 1, xl[1] - xl[0], 0, 1, yl[1] - yl[0], dz=0)
    vol = 1.0 / (numx * numy)
    p = _sys_from_lattice(x, vol)
    rng = np.random.default_rng(seed)
    u0 = scatter(rng, x.shape[0], count).reshape(-1, 1)
    nt, steps = timespan_steps(t_end, num_points, dt)
    sa = SystemArrays(**p, n_types=1, dimension=detect_dimension(x), h=find_h(x), dt=dt, nt=nt,
                      output_steps=steps, static_domain=1,
                      bbox=bounding_box(x), u0=u0.astype(np.uint32),
                      diffusion_matrix=np.array([[D]]),
                      reactions=[(K_ZERO, 0, 0, k_birth, 0, {}, {0: 1}),
                                 (K_FIRST, 0, 0, k_death, 0, {0: 1}, {})])
    sa.build_sparse()
    return sa.normalized()


def static_rd_3d(n_side=40, nA=100000, nB=100000, kf=1e-3, kr=0.1, DA=0.01, DB=0.01, DC=0.001,
                 t_end=100, num_points=101, dt=1.0, seed=0):
    """Config 2: static 3D reaction-diffusion A+B<->C on an n^3 lattice in
    the unit cube (create_3D_domain)."""
    l = np.linspace(0, 1, n_side)
    d = l[1] - l[0]
    x = cartesian_lattice(0, 1, d, 0, 1, d, 0, 1, d)
    vol = 1.0 / n_side ** 3
    p = _sys_from_lattice(x, vol)
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    u0 = np.stack([scatter(rng, n, nA), scatter(rng, n, nB), np.zeros(n, np.int64)], axis=1)
    nt, steps = timespan_steps(t_end, num_points, dt)
    sa = SystemArrays(**p, n_types=1, dimension=detect_dimension(x), h=find_h(x), dt=dt, nt=nt,
                      output_steps=steps, static_domain=1, bbox=bounding_box(x),
                      u0=u0.astype(np.uint32), diffusion_matrix=np.array([[DA], [DB], [DC]]),
                      reactions=[(K_AB, 0, 1, kf, 0, {0: 1, 1: 1}, {2: 1}),
                                 (K_FIRST, 2, 0, kr, 0, {2: 1}, {0: 1, 1: 1})])
    sa.build_sparse()
    return sa.normalized()


def dam_break_2d(nx=200, ny=200, spacing=0.0025, box=(2.0, 1.5), n_mol=1000, D=1e-3, k=0.05,
                 dt=1e-4, nt=10000, out_every=100, rho0=1.0, c0=10.0, nu=1e-3, seed=0):
    """Config 3: 2D SDPD dam break with transport.  Fluid nx x ny at
    `spacing` from (spacing, spacing); a 2-particle-thick fixed wall ring
    around [0,box[0]]x[0,box[1]] (the two lattice rows at and outside the
    box edges)."""
    s = spacing
    fi = np.arange(nx)
    fj = np.arange(ny)
    fx = np.array([((1 + i) * s, (1 + j) * s, 0.0) for i in fi for j in fj])
    nbx = int(round(box[0] / s))
    nby = int(round(box[1] / s))
    wall = []
    for i in range(-1, nbx + 2):
        for j in range(-1, nby + 2):
            if i <= 0 or i >= nbx or j <= 0 or j >= nby:
                wall.append((i * s, j * s, 0.0))
    wx = np.array(wall)
    x = np.concatenate([fx, wx])
    nf, nw = fx.shape[0], wx.shape[0]
    n = nf + nw
    mass = rho0 * s * s
    sa = SystemArrays(x=x, type=np.concatenate([np.ones(nf), 2 * np.ones(nw)]).astype(np.int32),
                      mass=np.full(n, mass), rho=np.full(n, rho0), nu=np.full(n, nu),
                      fixed=np.concatenate([np.zeros(nf), np.ones(nw)]).astype(np.int32),
                      n_types=2, dimension=2, h=1.3 * s, dt=dt, nt=nt,
                      output_steps=np.arange(0, nt + 1, out_every, dtype=np.uint32),
                      static_domain=0, rho0=rho0, c0=c0, P0=rho0 * c0 * c0,
                      gravity=(0.0, -1.0, 0.0),
                      bbox=(-s, box[0] + s, -s, box[1] + s, 0, 0))
    rng = np.random.default_rng(seed)
    xs = fx[:, 0]
    left = np.nonzero(xs <= np.quantile(xs, 0.10))[0]
    u0 = np.zeros((n, 1), np.int64)
    u0[:, 0] = scatter(rng, n, n_mol, allowed=left)
    sa.u0 = u0.astype(np.uint32)
    sa.diffusion_matrix = np.array([[D, D]])
    sa.reactions = [(K_FIRST, 0, 0, k, 0, {0: 1}, {})]
    sa.build_sparse()
    return sa.normalized()


# Time step of the 1M-particle cylinder.  The acoustic limit is
# 0.25 h / c0 = 3.8e-4 at h = 1.3 x 0.023249; at 2.5e-4 the open cylinder
# turns non-finite at step 21 (reference engine, oracle and this engine
# alike), while at 1.25e-4 the 1M run and the oracle's 20k-particle run at
# the same Courant number (dt scaled with the spacing) both stay finite over
# 2,000 steps.
CYLINDER_DT_1M = 1.25e-4


def cylinder_dt(n_target, radius=1.0, length=4.0):
    """CYLINDER_DT_1M scaled with the lattice spacing to n_target particles
    (the same Courant number c0 dt / h at every size)."""
    s = (math.pi * radius * radius * length / n_target) ** (1.0 / 3.0)
    s1m = (math.pi * radius * radius * length / 1_000_000) ** (1.0 / 3.0)
    return CYLINDER_DT_1M * s / s1m


def cylinder_3d(n_target=1_000_000, radius=1.0, length=4.0, rho0=1.0, c0=20.0, nu=1e-2,
                inflow=1.0, nA=None, nB=None, k=1e-2, D=1e-3, dt=None, nt=2000,
                out_every=None, seed=0, lattice_spacing=None, walls=0, h_factor=1.3):
    """Config 5: fluid filling a cylinder (axis x) on a cubic lattice with
    spacing (pi r^2 L / n_target)^(1/3), inflow v=(inflow,0,0) on the slab
    x < 2*spacing, species A+B->C.  walls > 0 adds that many lattice layers
    of fixed wall particles (type 2) around the fluid, radius in
    (radius, radius + walls*spacing], so the fluid is a pipe flow rather than
    a free jet (the SDPD's background-pressure transport term pushes a free
    surface outward).  dt=None: cylinder_dt(n_target)."""
    if dt is None:
        dt = cylinder_dt(n_target, radius, length)
    vol_total = math.pi * radius * radius * length
    s = lattice_spacing or (vol_total / n_target) ** (1.0 / 3.0)
    nxs = int(math.floor(length / s)) + 1
    rw = radius + walls * s
    nr = int(math.floor(rw / s))
    ys = np.arange(-nr, nr + 1) * s
    Y, Z = np.meshgrid(ys, ys, indexing="ij")
    r2 = Y * Y + Z * Z
    fluid = r2 <= radius * radius
    wall = (~fluid) & (r2 <= rw * rw) if walls > 0 else np.zeros_like(fluid)
    yz = np.stack([Y[fluid], Z[fluid]], axis=1)
    yzw = np.stack([Y[wall], Z[wall]], axis=1)
    xs = np.arange(nxs) * s
    xf = np.concatenate([np.column_stack([np.full(yz.shape[0], xv), yz]) for xv in xs])
    xw = (np.concatenate([np.column_stack([np.full(yzw.shape[0], xv), yzw]) for xv in xs])
          if yzw.shape[0] else np.zeros((0, 3)))
    x = np.concatenate([xf, xw])
    nf, n = xf.shape[0], xf.shape[0] + xw.shape[0]
    typ = np.concatenate([np.ones(nf, np.int32), np.full(n - nf, 2, np.int32)])
    fixed = (typ == 2).astype(np.int32)
    mass = rho0 * s ** 3
    nA = nf if nA is None else nA
    nB = nf if nB is None else nB
    rng = np.random.default_rng(seed)
    u0 = np.zeros((n, 3), np.int64)
    u0[:nf] = np.stack([scatter(rng, nf, nA), scatter(rng, nf, nB), np.zeros(nf, np.int64)], axis=1)
    out_every = out_every or nt
    nty = 2 if walls > 0 else 1
    sa = SystemArrays(x=x, type=typ, mass=np.full(n, mass), rho=np.full(n, rho0),
                      nu=np.full(n, nu), fixed=fixed, n_types=nty, dimension=3,
                      h=h_factor * s, dt=dt, nt=nt,
                      output_steps=np.arange(0, nt + 1, out_every, dtype=np.uint32),
                      static_domain=0, rho0=rho0, c0=c0, P0=rho0 * c0 * c0,
                      bbox=(0, length, -rw, rw, -rw, rw),
                      u0=u0.astype(np.uint32), diffusion_matrix=np.array([[D, 0.0][:nty]] * 3),
                      reactions=[(K_AB, 0, 1, k, 0, {0: 1, 1: 1}, {2: 1})],
                      bcs=[dict(lo=(-math.inf, -math.inf, -math.inf), hi=(2 * s, math.inf, math.inf),
                                type=1, target=1, value=(inflow, 0.0, 0.0))])
    sa.build_sparse()
    return sa.normalized()


def jittered_lattice_3d(n_target, jitter=0.1, seed=0):
    """Config 4 domain: a cubic lattice in the unit cube with about n_target
    points, each jittered by U(-jitter, jitter) x spacing per axis."""
    rng = np.random.default_rng(seed)
    m = max(2, int(round(n_target ** (1.0 / 3.0))))
    l = np.linspace(0, 1, m)
    s = l[1] - l[0]
    X, Y, Z = np.meshgrid(l, l, l, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    x = x + rng.uniform(-jitter * s, jitter * s, size=x.shape)
    return x, 1.0 / x.shape[0]


def gene_network_3d(n_target=8000, seed=0, t_end=200, dt=1.0):
    """Config 4 trajectory: 5-species negative-feedback gene network with 6
    mass-action reactions on a jittered 3D lattice; rates log-uniform in
    [1e-3, 10], D log-uniform in [1e-4, 1e-2], Poisson(50) initial counts per
    species per trajectory, scattered."""
    rng = np.random.default_rng(seed)
    x, vol = jittered_lattice_3d(n_target, seed=seed)
    n = x.shape[0]
    p = _sys_from_lattice(x, vol)
    lr = lambda lo, hi: float(10 ** rng.uniform(math.log10(lo), math.log10(hi)))
    G_on, G_off, M, P, Rp = 0, 1, 2, 3, 4
    rxns = [
        (K_FIRST, G_on, 0, lr(1e-3, 10), 0, {G_on: 1}, {G_on: 1, M: 1}),      # transcription
        (K_FIRST, M, 0, lr(1e-3, 10), 0, {M: 1}, {M: 1, P: 1}),               # translation
        (K_FIRST, P, 0, lr(1e-3, 10), 0, {P: 1}, {Rp: 1}),                     # maturation
        (K_AB, G_on, Rp, lr(1e-3, 10), 0, {G_on: 1, Rp: 1}, {G_off: 1}),       # repression
        (K_FIRST, G_off, 0, lr(1e-3, 10), 0, {G_off: 1}, {G_on: 1, Rp: 1}),    # release
        (K_FIRST, M, 0, lr(1e-3, 10), 0, {M: 1}, {}),                          # decay
    ]
    Ds = np.array([[lr(1e-4, 1e-2)] for _ in range(5)])
    u0 = np.stack([scatter(rng, n, int(rng.poisson(50))) for _ in range(5)], axis=1)
    nt, steps = timespan_steps(t_end, t_end + 1, dt)
    sa = SystemArrays(**p, n_types=1, dimension=3, h=find_h(x), dt=dt, nt=nt, output_steps=steps,
                      static_domain=1, bbox=(0, 1, 0, 1, 0, 1), u0=u0.astype(np.uint32),
                      diffusion_matrix=Ds, reactions=rxns)
    sa.build_sparse()
    return sa.normalized()


def custom_chemistry_2d(numx=12, numy=12, t_end=2, num_points=5, dt=0.1, seed=0, libm=False):
    """A static 2D model built through the Model API whose reactions are
    custom propensity functions (Reaction(propensity_function=...)): integer
    arithmetic of the generated C++ (A*(A-1)/2, B/3), a data function, type
    restrictions, sqrt, and with libm=True time dependence and pow/exp.  Two
    particle types (x > 0.5 is type 2)."""
    from .core import model as M
    from .core.domain import Domain
    from .solver import Solver

    class Gradient(M.DataFunction):
        def map(self, point):
            return 1.0 + 2.0 * point[0]

    m = M.Model("custom_chemistry")
    d = Domain.create_2D_domain(xlim=(0, 1), ylim=(0, 1), numx=numx, numy=numy, type_id=1, fixed=True)
    d.type_id[d.vertices[:, 0] > 0.5] = "type_2"
    m.add_domain(d)
    m.add_species([M.Species("A", 0.05), M.Species("B", 0.02), M.Species("C", 0.01, restrict_to=[1])])
    m.add_parameter([M.Parameter("k1", expression=40.0), M.Parameter("k2", expression=0.5),
                     M.Parameter("K", expression=6.0), M.Parameter("kd", expression=0.05)])
    m.add_data_function(Gradient(name="grad"))
    birth = "k1*vol*grad*exp(-t/4)" if libm else "k1*vol*grad"
    conv = "k2*A*pow(A,2)/(K*K+pow(A,2))" if libm else "k2*A/(1+A/K)"
    m.add_reaction([
        M.Reaction("birth", reactants={}, products={"A": 1}, propensity_function=birth),
        M.Reaction("conv", reactants={"A": 1}, products={"B": 1}, propensity_function=conv),
        M.Reaction("dimer", reactants={"A": 2}, products={"C": 1}, propensity_function="A*(A-1)/2*kd",
                   restrict_to=[1]),
        M.Reaction("decay", reactants={"B": 1}, products={}, propensity_function="kd*B + B/3"),
        M.Reaction("back", reactants={"C": 1}, products={"A": 1}, rate="kd"),
        M.Reaction("recycle", reactants={"B": 1}, products={"A": 1}, propensity_function="sqrt(B)*k2",
                   restrict_to=[2]),
    ])
    m.add_initial_condition(M.ScatterInitialCondition("A", 300, seed=seed))
    m.add_initial_condition(M.ScatterInitialCondition("B", 100, seed=seed + 1))
    m.timespan(M.TimeSpan.linspace(t=t_end, num_points=num_points, timestep_size=dt))
    return Solver(m).lower()
