"""The surrogate network's GPU kernels against the CPU oracle
(oracle/mpn_oracle.py, numpy FP64 with bf16-rounded matrix-product
operands, whose own backward is pinned by finite differences on the CPU).

Tolerances: the GPU rounds the same operands to bf16 but feeds them from
FP32 intermediates, so an operand occasionally rounds to the neighbouring
bf16 value; outputs agree to ~1e-3 relative and gradients to ~1e-2 (L2
over each parameter block), with cosine similarity > 0.999 overall."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mo():
    from oracle import mpn_oracle
    return mpn_oracle


def _setup(mo, n=300, seed=4, h=0.12):
    from spatialpy_amd.surrogate import Surrogate
    s = Surrogate(seed=1)
    assert s.P == mo.P
    w = mo.init_params(seed).astype(np.float32)
    s.set_params(w)
    x, X0, deg, nidx, inv_h, y = mo.random_graph(n=n, seed=seed, h=h)
    idx = s.add_sample(x, X0, deg, nidx, inv_h, y)
    return s, w.astype(np.float64), (x, X0, deg, nidx, inv_h, y), idx


def test_mpn_forward_matches_oracle(mo):
    s, w, (x, X0, deg, nidx, inv_h, y), idx = _setup(mo)
    out = s.forward(idx, X0.shape[0])
    ref, _ = mo.forward(w, x, X0, deg, nidx, inv_h, emulate_bf16=True)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 5e-3, rel
    # and far from an unrelated network: the check has teeth
    ref2, _ = mo.forward(mo.init_params(99), x, X0, deg, nidx, inv_h, emulate_bf16=True)
    assert np.linalg.norm(out - ref2) / np.linalg.norm(ref2) > 0.2


def test_mpn_gradient_matches_oracle(mo):
    s, w, (x, X0, deg, nidx, inv_h, y), idx = _setup(mo)
    loss, g = s.grad(idx)
    rl, rg = mo.loss_grad(w, x, X0, deg, nidx, inv_h, y, emulate_bf16=True)
    assert abs(loss - rl) / rl < 5e-3
    cos = float(g @ rg / (np.linalg.norm(g) * np.linalg.norm(rg)))
    assert cos > 0.999, cos
    gb, rb = mo.unpack(g), mo.unpack(rg)
    blocks = [("We", gb["We"], rb["We"]), ("be", gb["be"], rb["be"]), ("Wd", gb["Wd"], rb["Wd"]),
              ("bd", gb["bd"], rb["bd"])]
    for l in range(mo.LAYERS):
        for k in gb["layers"][l]:
            blocks.append((f"L{l}.{k}", gb["layers"][l][k], rb["layers"][l][k]))
    for name, a, b in blocks:
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 3e-2, (name, rel)


def test_mpn_training_reduces_loss(mo):
    s, w, (x, X0, deg, nidx, inv_h, y), idx = _setup(mo, n=400, seed=5)
    # a learnable target: a smooth function of the node features and of the
    # neighbour mean
    tgt = np.tanh(X0[:, :3] * 0.7).astype(np.float32)
    s.clear_samples()
    k = s.add_sample(x, X0, deg, nidx, inv_h, tgt)
    l0, _ = s.grad(k)
    l1 = s.train(epochs=300, lr=2e-3, seed=0)
    assert l1 < 0.1 * l0, (l0, l1)


def _dam_break(nt):
    from spatialpy_amd import workloads
    return workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=nt, out_every=nt, dt=2e-4)


def _trained(sa, warm=30, record_steps=200, every=4, epochs=60):
    from spatialpy_amd.engine import Engine
    from spatialpy_amd.surrogate import Surrogate, fit_scales
    eng = Engine(sa)
    eng.reset(seed=1)
    eng.advance(warm)
    s = Surrogate(seed=0)
    s.set_scales(*fit_scales(eng, sa))
    s.record(eng, every)
    eng.advance(record_steps)
    s.stop_recording(eng)
    assert s.num_samples() == record_steps // every
    return eng, s


def test_surrogate_training_gains_over_untrained_and_fallback_works():
    """Trained on plain GPU SDPD steps recorded inside the step, the network
    predicts the pair accelerations of later steps far better than the same
    network untrained (the method without training); with the 2% budget the
    first check falls back to the plain SDPD; with a loose budget the
    surrogate keeps running and the state stays finite."""
    from spatialpy_amd.surrogate import Surrogate
    sa = _dam_break(600)
    eng, s = _trained(sa)
    # untrained error at a check step
    s.attach(eng, check_every=1, max_rel_err=1e9)
    eng.advance(1)
    e0 = Surrogate.stats(eng)["last_err"]
    s.detach(eng)
    loss = s.train(60, lr=2e-3, seed=0)
    s.attach(eng, check_every=1, max_rel_err=1e9)
    eng.advance(1)
    e1 = Surrogate.stats(eng)["last_err"]
    print("relative error untrained", e0, "trained", e1, "loss", loss)
    assert e1 < 0.6 * e0 and e1 < 1.0
    # the 2% budget: the first check switches the network off
    s.attach(eng, check_every=5, max_rel_err=0.02)
    eng.advance(10)
    st = Surrogate.stats(eng)
    assert st["checks"] == 1 and not st["active"] and st["nn_steps"] >= 1
    # a loose budget keeps it on between checks
    s.attach(eng, check_every=5, max_rel_err=1e9)
    eng.advance(20)
    st = Surrogate.stats(eng)
    assert st["active"] and st["nn_steps"] >= 16 and st["checks"] >= 4
    x = eng.state()["x"]
    assert np.isfinite(x).all()
    # reactions and diffusion stay the exact SSA: the run's event counters
    # keep advancing with the network attached
    assert eng.stats().diffusions + eng.stats().reactions > 0
