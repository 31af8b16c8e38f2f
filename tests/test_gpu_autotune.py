"""Launch-shape autotuner on the GPU: the Robertson variants with 512, 256
and 128 registers per lane (prebuilt by __graft_entry__.build) are timed on a
2048-instance forward-sensitivity batch; every variant gives the oracle's
results, and the choice is the fastest measured and is reused from the cache."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from amici_amd import networks

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"
TIGHT = dict(rtol=1e-10, atol=1e-14, maxsteps=200000)


def test_autotune_robertson(gpu, tmp_path, monkeypatch):
    import torch

    from amici_amd import autotune, sim

    monkeypatch.setattr(autotune, "cache_path", lambda n: tmp_path / "autotune.json")
    net = networks.robertson()
    g = dict(np.load(GOLD / "robertson_batch64.npz"))
    rep = 32  # 2048 instances: more than one per SIMD
    theta, kfix = np.tile(g["theta"], (rep, 1)), np.tile(g["kfix"], (rep, 1))
    meas, sigma = np.tile(g["meas"], (rep, 1, 1)), np.tile(g["sigma"], (rep, 1, 1))
    solver = sim.Solver(**TIGHT)
    solver.set_sensitivity_order(sim.SensitivityOrder.first)
    solver.set_sensitivity_method(sim.SensitivityMethod.forward)
    r = autotune.tune(net, solver, theta, kfix, g["ts"], meas, sigma, reps=2)
    print("autotune", r.times_ms, "->", r.team, r.waves, flush=True)
    assert set(r.times_ms) == {"t64_w1", "t64_w2", "t64_w4"}
    assert all(v > 0 for v in r.times_ms.values())
    assert r.times_ms[autotune.label(dict(team=r.team, waves=r.waves))] == min(r.times_ms.values())
    assert not r.cached
    again = autotune.tune(net, solver, theta, kfix, g["ts"], meas, sigma,
                          loader=lambda c: pytest.fail("timed again"))
    assert again.cached and (again.team, again.waves) == (r.team, r.waves)
    # every variant computes the oracle's gradient
    for c in autotune.candidates(net.nx):
        m = sim.Model(net, team=c["team"], waves=c["waves"])
        out = sim.simulate_batch(m, solver, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"])
        torch.cuda.synchronize()
        assert (out["status"].cpu().numpy() == 0).all()
        sllh = out["sllh"].cpu().numpy()
        err = np.abs(sllh - g["sllh_fsa"]) / (1e-5 * np.abs(g["sllh_fsa"]) + 1e-7 * np.abs(
            g["sllh_fsa"]).max(axis=1, keepdims=True))
        assert err.max() <= 1.0, (autotune.label(c), err.max())
