"""The learned warm start end to end on the GPU (pecs_amd.warmstart).

Steady states of 256 seeded draws on refinements 3 and 4 (configuration
physics, k = 2; cone angle, applied bias, generation amplitude and doping
drawn) train the encoder + convolutional decoder on the GPU; on 32 held-out
draws the network's 64 x 64 grid predictions are projected onto each
instance's own mesh and stepped to the relative steady residual 1e-8.
Asserted: every instance converges; exactly the instances whose predicted
residual (the relative density change over the first step, about half the
prediction's relative error) exceeds 1e-2 fall back, at most half of them
with this small training set; the predictions are within 5 % (median) of
the steady states; the accepted instances need at least 15 % fewer steps
than the cold start from the equilibrium state on the same draws, and all
instances together fewer than the cold start.  A deliberately wrong
prediction (log-normal noise on every grid value) is caught by the check
and restarts cold, reaching the steady state in the cold start's steps plus
the probe step.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

K, DT, TOL, EVERY, MAXS = 2, 1e-3, 1e-8, 25, 20000


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def trained():
    from pecs_amd import warmstart as ws
    rng = np.random.default_rng(2024)
    train_groups = ws.draw_groups(rng, 16, 16, refines=(3, 4))
    test_groups = ws.draw_groups(rng, 2, 16, refines=(3, 4))
    _say(f"[gpu] warm start: generating 256 steady states on refinements {[g[0] for g in train_groups]}")
    X, Y, steps, wall, setup = ws.generate(train_groups, K, DT, TOL, EVERY, MAXS)
    assert np.all(steps > 0)
    _say(f"[gpu] warm start: training ({wall:.1f} s stepping, cold steps median {np.median(steps)})")
    net, norm, hist = ws.train(X, Y, epochs=600, batch=32)
    _say(f"[gpu] warm start: loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    assert hist[-1] < 0.05 * hist[0]
    yield ws, net, norm, test_groups
    net.close()


def test_warm_start_beats_cold_start(trained):
    ws, net, norm, test_groups = trained
    _say(f"[gpu] warm start: cold start of 32 held-out draws on refinements {[g[0] for g in test_groups]}")
    cold, wall_cold, steady = ws.cold_start(test_groups, K, DT, TOL, EVERY, MAXS)
    _say("[gpu] warm start: warm start of the same draws")
    ev = ws.evaluate(net, norm, test_groups, K, DT, TOL, EVERY, MAXS)
    warm = ev["steps"]
    _say(f"  steps cold mean {cold.mean():.0f}, warm mean {warm.mean():.0f}; wall cold {wall_cold:.2f} s, "
         f"warm {ev['wall']:.2f} s + {ev['predict_s']:.3f} s predict/project; predicted residual "
         f"median {np.median(ev['predicted_residual']):.2e} max {ev['predicted_residual'].max():.2e}")
    assert np.all(cold > 0) and np.all(warm > 0)
    rho = ev["predicted_residual"]
    fell = np.zeros(len(warm), bool)
    ninst = [len(g[2]) for g in test_groups]
    for ci, i in ev["fell"]:
        fell[sum(ninst[:ci]) + i] = True
    assert np.array_equal(fell, rho > 1e-2)
    assert fell.sum() <= len(warm) // 2
    kept = ~fell
    _say(f"  accepted {kept.sum()}: warm mean {warm[kept].mean():.0f} against cold {cold[kept].mean():.0f}")
    assert warm[kept].mean() <= 0.85 * cold[kept].mean()
    assert warm.mean() < cold.mean()
    # the predictions against the steady states of the cold start
    errs = [np.linalg.norm(p - y) / np.linalg.norm(y) for p, y in zip(ev["pred"], steady)]
    _say(f"  prediction error (relative L2 on the grid): median {np.median(errs):.2e} max {np.max(errs):.2e}")
    assert np.median(errs) < 0.05


def test_wrong_prediction_takes_the_fallback(trained):
    ws, net, norm, test_groups = trained
    groups = test_groups[:1]
    cold, _, _ = ws.cold_start(groups, K, DT, TOL, EVERY, MAXS)
    rng = np.random.default_rng(5)
    ev = ws.evaluate(net, norm, groups, K, DT, TOL, EVERY, MAXS,
                     perturb=lambda p: p * np.exp(0.7 * rng.standard_normal(p.shape)))
    _say(f"[gpu] warm start: wrong predictions: predicted residual min {ev['predicted_residual'].min():.2e}, "
         f"fallbacks {ev['fallbacks']}")
    assert ev["fallbacks"] == len(cold)
    assert np.all(ev["predicted_residual"] > 1e-2)
    assert np.all(np.abs(ev["steps"] - 1 - cold) <= EVERY), (ev["steps"], cold)
