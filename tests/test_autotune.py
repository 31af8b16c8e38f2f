"""Launch-shape autotuner, host side: candidate shapes, variant modules and
the choice and its cache (timings injected; the GPU test times real kernels)."""
from __future__ import annotations

import json

import numpy as np
import pytest

from amici_amd import autotune, codegen, networks, sim


def test_candidates_by_state_count():
    assert [autotune.label(c) for c in autotune.candidates(3)] == ["t64_w1", "t64_w2", "t64_w4"]
    assert [autotune.label(c) for c in autotune.candidates(40)] == ["t64_w1", "t64_w2"]
    assert [autotune.label(c) for c in autotune.candidates(100)] == ["t128_w1", "t256_w1"]
    assert [autotune.label(c) for c in autotune.candidates(200)] == ["t256_w1"]


def test_variant_sources_and_names():
    net = networks.robertson()
    assert "#define AMX_WAVES_PER_EU 4" in codegen.emit_model_source(net, waves=4)
    assert "AMX_WAVES_PER_EU" not in codegen.emit_model_source(net)
    assert codegen.variant_name(net) == "robertson"
    assert codegen.variant_name(net, waves=2) == "robertson"  # the default budget
    assert codegen.variant_name(net, waves=4) == "robertson_w4"
    n200 = networks.network_n200()
    assert codegen.variant_name(n200, team=256) == n200.name
    with pytest.raises(ValueError):
        codegen.emit_model_source(net, waves=3)
    with pytest.raises(ValueError):
        codegen.emit_model_source(n200, team=128)  # fewer lanes than states


def test_tune_picks_fastest_and_caches(tmp_path, monkeypatch):
    net = networks.robertson()
    monkeypatch.setattr(autotune, "cache_path", lambda n: tmp_path / "autotune.json")
    ms = {"t64_w1": 3.0, "t64_w2": 1.5, "t64_w4": 2.0}
    timed = []

    def timer(model, solver, batch, reps, warmup):
        timed.append(model)
        assert batch["theta"].shape == (8, 3) and reps == 3 and warmup == 1
        return ms[model]

    solver = sim.Solver()
    theta = np.tile(net.theta_nominal, (8, 1))
    r = autotune.tune(net, solver, theta, np.zeros((8, 0)), np.linspace(1, 10, 5), device="dev",
                      timer=timer, loader=autotune.label)
    assert (r.team, r.waves, r.cached) == (64, 2, False) and r.times_ms == ms
    assert sorted(timed) == sorted(ms)
    cache = json.loads((tmp_path / "autotune.json").read_text())
    key = autotune.cache_key("dev", 8, 5, solver)
    assert cache[key]["waves"] == 2
    # the same shape again: from the cache, nothing timed
    r2 = autotune.tune(net, solver, theta, np.zeros((8, 0)), np.linspace(1, 10, 5), device="dev",
                       timer=None, loader=lambda c: pytest.fail("loaded a variant"))
    assert (r2.team, r2.waves, r2.cached) == (64, 2, True)
    # another solver setting is another entry
    solver.set_sensitivity_order(sim.SensitivityOrder.first)
    assert autotune.cache_key("dev", 8, 5, solver) != key
