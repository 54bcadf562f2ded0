"""Host-side pieces of the learned warm start and the bench contract that run
without a GPU: parameter draws and features, the output normaliser, the
decoder's output grid, and bench.py's refusal of more than one GPU."""
import os
import subprocess
import sys

import numpy as np

from pecs_amd import warmstart as ws

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_draws_and_features():
    rng = np.random.default_rng(1)
    groups = ws.draw_groups(rng, 50, 16, refines=(5, 6, 7, 8))
    assert len(groups) == 50 and all(len(inst) == 16 for _, _, inst in groups)
    refs = np.array([r for r, _, _ in groups])
    assert set(refs.tolist()) <= {5, 6, 7, 8} and len(set(refs.tolist())) > 1
    angs = np.array([a for _, a, _ in groups])
    assert angs.min() >= 0.0 and angs.max() <= 15.0
    inst = np.array([d for _, _, i in groups for d in i])
    app, g0, ne = inst[:, 0], inst[:, 1], inst[:, 2]
    assert np.all(np.abs(app) <= 0.5) and np.all((g0 >= 0.1) & (g0 <= 10)) and np.all((ne >= 0.5) & (ne <= 5))
    f = np.array([ws.features(a, *d) for _, a, i in groups for d in i])
    assert f.shape == (800, 4) and np.all(np.abs(f) <= 1.01)   # "about [-1, 1]"
    p = ws.instance_params(5, 2, 5e-4, 10.0, 0.2, 1.0, 2.0)
    assert np.isclose(p["r2"] - p["r1"], np.tan(np.radians(10.0))) and p["L"] == 1.5 and p["phi_app"] == 0.2


def test_normaliser_round_trip():
    rng = np.random.default_rng(2)
    Y = rng.normal(size=(40, 7)) * [1, 1e3, 1e-3, 5, 0, 2, 1] + [0, 5, 1, 0, 3, 0, 0]
    n = ws.Normaliser(Y)
    Z = n.fwd(Y)
    assert Z.dtype == np.float32
    assert np.allclose(Z[:, [0, 1, 3, 5, 6]].mean(axis=0), 0, atol=1e-5)
    assert np.allclose(n.back(Z), Y, rtol=1e-6, atol=1e-6 * np.abs(Y).max())
    assert np.all(np.isfinite(Z[:, 4]))   # a constant column does not divide by zero


def test_decoder_shape():
    # the decoder's 64 x 64 grid of the four density channels
    s = ws.DECODER
    G = s["g0"] * 2 ** sum(u for _, u in s["stages"])
    assert G == ws.GRID == 64 and s["stages"][-1][0] == 4


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "one GPU" in r.stdout
