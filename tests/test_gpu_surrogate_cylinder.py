"""The surrogate on the cylinder workload (A+B->C carried by a 3D SDPD free
jet with inflow), at 64k particles over the configured 2,000 steps:
trained on the plain GPU SDPD steps of that run (a sample of 8,192 moving
particles every 25 steps, 2,000 epochs: the ensemble shift below halves from
800 epochs to 2,000, and does not follow the online-refresh settings),
attached with a check against the plain pair kernel every 50 steps and the
2% budget, re-armed (a failed check pauses the network until the next
check passes) and refreshed online from every check's plain step (the
surrogate-driven flow drifts from the plain run it was trained on).
Training sums its gradients in a fixed order, so the run is reproducible.

Asserted: every check runs; only a few fail (each pauses the network for
one interval, so every interval the network drives starts from a check
within the 2% budget); the network drives at least 80% of the steps; and
the ensemble-mean species totals of 32 surrogate trajectories fall inside
the plain method's 95% band (mean +- 1.96 standard deviations of 32 plain
trajectories) at every output time.  The two-sample z of the means is
printed beside it with its family-wise 95% level (Bonferroni over output
times and species): with SSA noise alone between the ensembles, a shift of
a tenth of a percent in a count shows there.  Reactions and diffusion are exact SSA in both
ensembles; only the particle motion differs."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N_TARGET = 64_000
STEPS = 2000


def _z(m):
    """Two-sided normal quantile for a 95% family over m comparisons."""
    from scipy.stats import norm
    return float(norm.ppf(1.0 - 0.05 / (2.0 * m)))


def test_cylinder_surrogate_stays_in_budget_and_in_the_species_band():
    from spatialpy_amd import workloads
    from spatialpy_amd.engine import Engine
    from spatialpy_amd.surrogate import Surrogate, fit_scales
    sa = workloads.cylinder_3d(n_target=N_TARGET, nt=STEPS + 1, out_every=250)
    print(f"cylinder: n={sa.n} dt={sa.dt:g} steps={STEPS}", flush=True)
    # scales from the state at mid-run, then samples over the whole plain run
    eng = Engine(sa)
    eng.reset(seed=1)
    eng.advance(STEPS // 2)
    s = Surrogate(seed=0)
    s.set_scales(*fit_scales(eng, sa))
    eng.reset(seed=1)
    s.record(eng, 25, max_particles=8192)
    eng.advance(STEPS)
    s.stop_recording(eng)
    eng.close()
    assert s.num_samples() == STEPS // 25 - 1
    loss = s.fit(2000, lr=3e-3, lr_end=3e-5)
    train_err = s.relative_error()
    print(f"trained on {s.num_samples()} samples: loss {loss:.3g}, training-set error {train_err:.4f}", flush=True)
    assert train_err < 0.01

    B = 32
    plain = Engine(sa)
    res_p = plain.run_batch(np.arange(B) + 100, fields=())
    plain.close()
    nn = Engine(sa)
    # online refresh: each check's plain step becomes a sample and the
    # network takes 8 Adam steps (half of them on that sample)
    s.attach(nn, check_every=50, max_rel_err=0.02, refresh_iters=8, refresh_lr=1e-4, keep_samples=400,
             max_particles=8192, rearm=True)
    res_n = nn.run_batch(np.arange(B) + 900, fields=())
    st = Surrogate.stats(nn)
    print(f"surrogate run: {st}", flush=True)
    nn.close()
    assert st["checks"] == STEPS // 50 and st["refreshes"] == st["checks"], st
    assert st["failed_checks"] <= 4, st
    assert 0 < st["max_err"] <= 0.02, st
    assert st["nn_steps"] >= 0.8 * STEPS, st

    tp = res_p["xx"].astype(np.int64).sum(axis=2)      # [B][K][S]
    tn = res_n["xx"].astype(np.int64).sum(axis=2)
    assert np.array_equal(tp[:, 0], tn[:, 0])          # both start from u0
    K, S = tp.shape[1], tp.shape[2]
    zf = _z((K - 1) * S)
    for k in range(1, K):
        for sp in range(S):
            mp, mn = tp[:, k, sp].mean(), tn[:, k, sp].mean()
            sd = tp[:, k, sp].std(ddof=1)
            se = math.sqrt(tp[:, k, sp].var(ddof=1) / B + tn[:, k, sp].var(ddof=1) / B)
            print(f"snapshot {k} species {sp}: plain {mp:.1f} +- {1.96 * sd:.1f} (95% band), surrogate mean "
                  f"{mn:.1f} ({(mn - mp) / max(sd, 1e-9):+.2f} sd); difference of means "
                  f"{abs(mn - mp) / max(se, 1e-9):.2f} standard errors "
                  f"(family-wise 95% two-sample level {zf:.2f})", flush=True)
            # the plain method's 95% band: where 95% of its trajectories fall
            assert abs(mn - mp) <= 1.96 * sd + 1e-9, (k, sp, mp, mn, sd)
    # the band has teeth: the C total moves by many band widths over the run,
    # so a surrogate that froze the chemistry would fall outside it
    assert tp[:, -1, 2].mean() - tp[:, 1, 2].mean() > 10 * 1.96 * tp[:, -1, 2].std(ddof=1)
