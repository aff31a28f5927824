"""Invariants and closed forms that pin the engine beyond the fixtures.

* Pair forces (model.cpp:39-191) on a lattice at rest at rho0 with no
  gravity: every term is exactly zero (pressure P0*(rho/rho0-1) = 0, dv = 0,
  vt - v = 0), so F and Frho are bitwise 0 -- on the oracle (CPU) and on
  both GPU force kernels; a particle without neighbours feels gravity only.
* Birth-death with the textbook selection (FLAG_TEXTBOOK_SELECTION): the
  species total is an immigration-death process, N(t) = Binomial(N0, e^-mu t)
  + Poisson(lambda/mu (1 - e^-mu t)) with lambda = k_birth * sum(vol) and
  mu = k_death, so its mean and variance at every output time are known in
  closed form; ensembles of the oracle and of the GPU's seeded stream fall
  inside their 99.9% bands."""
import numpy as np
import pytest
from scipy import stats

from spatialpy_amd import _abi, workloads
from spatialpy_amd.lowering import SystemArrays


def _lattice_at_rest(nx=12, ny=12, dx=0.1, lone=False):
    x = workloads.cartesian_lattice(0, (nx - 1) * dx, dx, 0, (ny - 1) * dx, dx)
    if lone:
        x = np.vstack([x, [[5.0, 5.0, 0.0]]])     # far from everything
    n = x.shape[0]
    vol = dx * dx
    sa = SystemArrays(x=x, type=np.ones(n, np.int32), mass=np.full(n, vol), rho=np.full(n, 1.0),
                      nu=np.full(n, 1e-2), fixed=np.zeros(n, np.int32), n_types=1, dimension=2,
                      h=2.2 * dx, dt=1e-4, nt=3, output_steps=np.array([0, 1, 2, 3]), static_domain=0,
                      rho0=1.0, c0=10.0, P0=100.0, gravity=(0.0, 0.0, 0.0),
                      bbox=(-1.0, 6.0, -1.0, 6.0, 0.0, 0.0))
    return sa.normalized()


def test_oracle_forces_vanish_on_a_lattice_at_rest(oracle):
    sa = _lattice_at_rest()
    o = oracle.Oracle(sa)
    o.reset(seed=1)
    o.advance(1)
    F, Fbp, Frho = o.eval_forces()
    assert not F.any() and not Frho.any()
    # the background-pressure term cancels only for interior particles
    inner = np.all((sa.x[:, :2] > 0.25) & (sa.x[:, :2] < 0.85), axis=1)
    assert np.abs(Fbp[inner]).max() < 1e-9 * np.abs(Fbp).max()


def _moments(sa, t):
    vol = sa.mass / sa.rho
    kb, kd = sa.reactions[0][3], sa.reactions[1][3]
    lam, n0 = kb * vol.sum(), int(sa.u0.sum())
    p = np.exp(-kd * t)
    mean = n0 * p + lam / kd * (1 - p)
    var = n0 * p * (1 - p) + lam / kd * (1 - p)
    return mean, var


def _check_moments(sa, totals, times):
    B = totals.shape[0]
    for k, t in enumerate(times):
        if t == 0:
            continue
        m, v = _moments(sa, t)
        a = totals[:, k].astype(float)
        z = (a.mean() - m) / np.sqrt(v / B)
        assert abs(z) < stats.norm.ppf(1 - 5e-4), (t, a.mean(), m)
        chi = a.var(ddof=1) * (B - 1) / v
        assert stats.chi2.ppf(5e-4, B - 1) < chi < stats.chi2.ppf(1 - 5e-4, B - 1), (t, a.var(ddof=1), v)


def _bd_textbook():
    sa = workloads.birth_death_2d(seed=1, t_end=10, num_points=6)
    sa.flags |= _abi.FLAG_TEXTBOOK_SELECTION
    return sa


def test_oracle_textbook_birth_death_moments_match_closed_form(oracle):
    sa = _bd_textbook()
    o = oracle.Oracle(sa)
    runs = [o.run(seed=s) for s in range(300)]
    tot = np.stack([r["xx"].sum(axis=(1, 2)) for r in runs])
    # snapshots are taken before steps 0, 1, 2, 4, ... (the reference's
    # output schedule), i.e. at t = step * dt
    times = runs[0]["steps"].astype(float) * sa.dt
    _check_moments(sa, tot, times)


@pytest.mark.gpu
def test_gpu_textbook_birth_death_moments_match_closed_form():
    from spatialpy_amd.engine import Engine
    sa = _bd_textbook()
    res = Engine(sa).run_batch(list(range(1, 401)), fields=("xx",))
    tot = res["xx"].sum(axis=(2, 3))
    times = res["steps"].astype(float) * sa.dt
    _check_moments(sa, tot, times)


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
def test_gpu_forces_vanish_on_a_lattice_at_rest_and_a_lone_particle_feels_gravity(fp32):
    from spatialpy_amd.engine import Engine
    sa = _lattice_at_rest(lone=True)
    eng = Engine(sa)
    eng.reset(seed=1)
    eng.advance(1)
    F, Fbp, Frho = eng.eval_forces(fp32=fp32)
    assert not F.any() and not Frho.any()
    sa2 = _lattice_at_rest(lone=True)
    sa2.gravity = (0.0, -1.0, 0.0)
    eng2 = Engine(sa2)
    eng2.reset(seed=1)
    eng2.advance(1)
    F2, _, _ = eng2.eval_forces(fp32=fp32)
    assert np.array_equal(F2[-1], [0.0, -1.0, 0.0])
