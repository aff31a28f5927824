"""The surrogate network's GPU kernels against the CPU oracle
(oracle/mpn_oracle.py: numpy FP64 with bf16 weights and bf16 hi + lo
activations, as the matrix-core products use them; its backward is pinned
by finite differences and its output space by the SDPD pair terms, both on
the CPU).

Tolerances: features are computed in FP64 and rounded to FP32 on both
sides; the GPU accumulates products in FP32, the oracle in FP64, so the
fused inference kernel and the training path agree with the oracle to
~1e-5 relative (asserted: 2e-4).  Gradients pass through bf16
weight-gradient products: ~1e-2 per parameter block (asserted 3e-2),
cosine similarity > 0.999 overall."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mo():
    from oracle import mpn_oracle
    return mpn_oracle


def _net(mo, sc, seed=4):
    from spatialpy_amd.surrogate import Surrogate
    s = Surrogate(seed=1)
    assert s.P == mo.P
    w = mo.init_params(seed).astype(np.float32)
    s.set_params(w)
    s.set_scales(*[float(x) for x in sc])
    return s, w.astype(np.float64)


def _oracle_acc(mo, w, g, sc, sel):
    ref, _ = mo.forward(w, g, sc, sel, emulate=True)
    ref = ref * np.float32(sc[4])
    ref[np.asarray(g["solid"])[sel] != 0] = 0.0
    ref[:, g["dim"]:] = 0.0
    return ref


@pytest.mark.parametrize("dim", [2, 3])
def test_fused_inference_matches_oracle(mo, dim):
    g = mo.random_graph(n=400, seed=4, h=0.12 if dim == 3 else 0.07, dim=dim)
    sc = mo.default_scales(g)
    s, w = _net(mo, sc)
    print(f"fused inference: dim={dim} n={g['n']} edges={int(g['deg'].sum())}", flush=True)
    out = s.predict(g)
    ref = _oracle_acc(mo, w, g, sc, np.arange(g["n"]))
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    assert rel < 2e-4, rel
    assert (out[g["solid"] != 0] == 0).all() and (out[:, dim:] == 0).all()
    # far from an unrelated network: the check has teeth
    ref2 = _oracle_acc(mo, mo.init_params(99), g, sc, np.arange(g["n"]))
    assert np.linalg.norm(out - ref2) / np.linalg.norm(ref2) > 0.2


def test_training_path_matches_oracle_and_fused_kernel(mo):
    g = mo.random_graph(n=400, seed=5, h=0.12)
    sc = mo.default_scales(g)
    s, w = _net(mo, sc)
    acc = np.random.default_rng(0).normal(0, 1, (g["n"], 3))
    k = s.add_sample(g, acc)
    sel = np.nonzero(g["solid"] == 0)[0]
    assert s.sample_size(k) == (len(sel), int(g["deg"][sel].sum()))
    pred, tgt = s.forward(k)
    np.testing.assert_allclose(tgt, acc[sel], rtol=1e-6, atol=1e-7)
    ref = _oracle_acc(mo, w, g, sc, sel)
    assert np.linalg.norm(pred - ref) / np.linalg.norm(ref) < 2e-4
    fused = s.predict(g)[sel]
    assert np.linalg.norm(pred - fused) / np.linalg.norm(fused) < 2e-4
    # a subset selection materialises only those particles' edges
    sub = sel[::3]
    k2 = s.add_sample(g, acc, sel=sub)
    p2, _ = s.forward(k2)
    assert s.sample_size(k2)[0] == len(sub)
    assert np.linalg.norm(p2 - fused[::3]) / np.linalg.norm(fused[::3]) < 2e-4


def test_gradient_matches_oracle(mo):
    g = mo.random_graph(n=300, seed=6, h=0.12)
    sc = mo.default_scales(g)
    s, w = _net(mo, sc)
    acc = np.random.default_rng(1).normal(0, 1, (g["n"], 3))
    k = s.add_sample(g, acc)
    sel = np.nonzero(g["solid"] == 0)[0]
    loss, gr = s.grad(k)
    y = (acc[sel].astype(np.float64) / np.float32(sc[4])).astype(np.float32)
    rl, rg = mo.loss_grad(w, g, sc, sel, y, emulate=True)
    assert abs(loss - rl) / rl < 1e-3, (loss, rl)
    cos = float(gr @ rg / (np.linalg.norm(gr) * np.linalg.norm(rg)))
    assert cos > 0.999, cos
    gb, rb = mo.unpack(gr), mo.unpack(rg)
    for name in gb:
        a, b = gb[name], rb[name]
        if name == "W1":
            a, b = a[:mo.NF], b[:mo.NF]
            assert (gb["W1"][mo.NF:] == 0).all()
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert rel < 3e-2, (name, rel)


def test_prediction_rotates_with_the_frame(mo):
    """Features are invariants and the outputs multiply vectors of the
    frame: rotating and reflecting every position, velocity and transport
    velocity rotates the predicted accelerations the same way."""
    g = mo.random_graph(n=400, seed=8, h=0.12)
    sc = mo.default_scales(g)
    s, _ = _net(mo, sc)
    out = s.predict(g)
    q, _ = np.linalg.qr(np.random.default_rng(3).normal(size=(3, 3)))
    q[:, 0] *= -1.0      # include a reflection
    g2 = dict(g)
    for k in ("x", "v", "vt"):
        g2[k] = np.ascontiguousarray(g[k] @ q.T)
    out2 = s.predict(g2)
    assert np.linalg.norm(out2 - out @ q.T) / np.linalg.norm(out) < 1e-4


def _dam_break(nt):
    from spatialpy_amd import workloads
    return workloads.dam_break_2d(nx=20, ny=20, spacing=0.025, n_mol=100, nt=nt, out_every=nt, dt=2e-4)


def test_surrogate_learns_the_dam_break_and_falls_back():
    """Trained on plain GPU SDPD steps recorded inside the step, the network
    predicts the pair accelerations of later steps to within 10% (an untrained
    one is off by ~100%); a budget below its error falls back at the
    first check; a budget above it keeps the network on between checks with
    finite state, and the event counters keep advancing (reactions and
    diffusion stay exact SSA)."""
    from spatialpy_amd.engine import Engine
    from spatialpy_amd.surrogate import Surrogate, fit_scales, relative_error
    sa = _dam_break(800)
    eng = Engine(sa)
    eng.reset(seed=1)
    eng.advance(30)
    s = Surrogate(seed=0)
    s.set_scales(*fit_scales(eng, sa))
    e_untrained = relative_error(eng, sa, s)
    s.record(eng, 4)
    eng.advance(240)
    s.stop_recording(eng)
    assert s.num_samples() == 60
    loss = s.fit(300, lr=3e-3, lr_end=3e-5)
    eng.advance(10)
    e = relative_error(eng, sa, s)
    print(f"dam break: relative error untrained {e_untrained:.4f} trained {e:.4f} (loss {loss:.3g}, "
          f"training-set error {s.relative_error():.4f})", flush=True)
    assert e < 0.1 and e < 0.1 * e_untrained
    # a budget below the error: steps 281..284 take the network's
    # accelerations, the check at step 285 switches it off
    eng.advance(1)
    s.attach(eng, check_every=5, max_rel_err=1e-6)
    eng.advance(10)
    st = Surrogate.stats(eng)
    assert st["checks"] == 1 and not st["active"] and st["nn_steps"] == 4, st
    # a loose budget keeps it on between checks
    s.attach(eng, check_every=5, max_rel_err=0.2)
    ev0 = eng.stats().diffusions + eng.stats().reactions
    eng.advance(40)
    st = Surrogate.stats(eng)
    assert st["active"] and st["nn_steps"] >= 32 and st["checks"] >= 8, st
    assert 0 < st["last_err"] < 0.2
    assert np.isfinite(eng.state()["x"]).all()
    assert eng.stats().diffusions + eng.stats().reactions > ev0


def test_training_is_reproducible(mo):
    """Gradients are summed in a fixed order (per-block partials added in
    block order, no float atomics): two networks with the same parameters
    trained on the same samples end bit-identical."""
    g = mo.random_graph(n=600, seed=9, h=0.12)
    sc = mo.default_scales(g)
    acc = np.random.default_rng(2).normal(0, 1, (g["n"], 3))
    outs = []
    for _ in range(2):
        s, _ = _net(mo, sc)
        s.add_sample(g, acc)
        s.add_sample(g, -acc, sel=np.arange(0, g["n"], 2))
        s.train(5, lr=1e-3, seed=3)
        outs.append(s.params())
    assert np.array_equal(outs[0], outs[1])
