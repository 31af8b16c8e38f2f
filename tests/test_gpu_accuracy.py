"""GPU accuracy at the benchmark's own tolerances, the mixed workload's
Robertson conditions, work-queue backfill and the per-output-time solver
diagnostics.

At the timed settings (rtol 1e-8, atol 1e-12) neither the GPU path nor CVODES
meets the rtol-1e-10 oracle to the parity bar of the tight-tolerance tests:
both carry the truncation error of rtol 1e-8.  The bar here is therefore
relative to the reference solver itself: the distribution of the GPU's errors
against the oracle at rtol 1e-10 over a batch (median, 90th percentile,
maximum, and each instance) may not exceed ten times that of CVODES run at the
same settings (fixtures n40_conditions16_rtol8 and robertson_mix1024, written
by tests/golden/make_golden.py), with the gradient measured as a vector
(max |error| / max |sllh|).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from amici_amd import networks

GOLD = Path(__file__).resolve().parent / "golden"
BENCH = dict(rtol=1e-8, atol=1e-12, quad_rtol=1e-8, quad_atol=1e-12, maxsteps=10000)
FACTOR = 10.0

pytestmark = pytest.mark.gpu


def _np(out):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
            if k != "_keepalive"}


def _solver(meth=None, **over):
    from amici_amd import sim
    sol = sim.Solver(**{**BENCH, **over})
    if meth is not None:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    return sol


def _normwise(a, ref):
    return np.abs(a - ref).max(axis=1) / np.abs(ref).max(axis=1)


def _entrywise(a, ref, floor):
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    scale = scale.reshape((-1,) + (1,) * (ref.ndim - 1))
    return (np.abs(a - ref) / (np.abs(ref) + floor * scale)).reshape(ref.shape[0], -1).max(axis=1)


def _gate(gpu_err, cv_err, what, cap=None):
    """The batch's error distribution against CVODES's at the same
    tolerances: the median, the 90th percentile and the largest GPU error
    within FACTOR x CVODES's, and each instance within FACTOR x the larger of
    CVODES's error for it and CVODES's 90th percentile.  (At rtol 1e-8 the
    error of one condition is itself a chaotic function of round-off: for a
    Robertson condition CVODES's own normwise gradient error ranges from
    2e-6 to 1.2e-4 under 1e-10 perturbations of log10 k, so a single CVODES
    run does not bound one instance's error.)"""
    q90 = np.quantile(cv_err, 0.9)
    bound = FACTOR * np.maximum(cv_err, q90)
    worst = int(np.argmax(gpu_err / bound))
    assert (gpu_err <= bound).all(), (
        f"{what}: instance {worst} error {gpu_err[worst]:.3g} > {FACTOR} x CVODES "
        f"{max(cv_err[worst], q90):.3g}")
    assert gpu_err.max() <= FACTOR * cv_err.max(), (what, gpu_err.max(), cv_err.max())
    assert np.quantile(gpu_err, 0.9) <= FACTOR * q90, (what, np.quantile(gpu_err, 0.9), q90)
    assert np.median(gpu_err) <= FACTOR * np.median(cv_err), (what, np.median(gpu_err))
    if cap is not None:
        assert gpu_err.max() <= cap, (what, gpu_err.max(), cap)


def test_n40_at_timed_tolerances(gpu):
    """The timed N40 adjoint settings on the 16 fixture conditions: states,
    observables, llh and sllh against the oracle at rtol 1e-10, gated by
    CVODES's own error at rtol 1e-8; the gradient also within 1e-5 (as a
    vector) of the oracle."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / "n40_conditions16.npz"))
    c = dict(np.load(GOLD / "n40_conditions16_rtol8.npz"))
    model = sim.Model(networks.network_n40())
    out = _np(sim.simulate_batch(model, _solver(sim.SensitivityMethod.adjoint), g["theta"],
                                 g["kfix"], g["ts"], g["meas"], g["sigma"]))
    torch.cuda.synchronize()
    assert (out["status"] == 0).all(), out["status"]
    _gate(_entrywise(out["x"], g["x"], 1e-9), _entrywise(c["x"], g["x"], 1e-9), "x")
    _gate(_entrywise(out["y"], g["y"], 1e-9), _entrywise(c["y"], g["y"], 1e-9), "y")
    _gate(np.abs(out["llh"] - g["llh"]) / np.abs(g["llh"]),
          np.abs(c["llh"] - g["llh"]) / np.abs(g["llh"]), "llh")
    _gate(_normwise(out["sllh"], g["sllh_asa"]), _normwise(c["sllh_asa"], g["sllh_asa"]),
          "sllh (adjoint, normwise)", cap=1e-5)


def test_robertson_mix_adjoint(gpu):
    """All 1024 Robertson conditions of the mixed workload, adjoint gradient
    at the benchmark's tolerances: every condition succeeds (as in CVODES),
    the gradient is gated by CVODES's own error (conditions where the oracle's
    rtol-1e-10 backward problem itself fails are compared with CVODES at rtol
    1e-8 instead), and the forward and backward step counts agree with
    CVODES's in total to 2 %."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / "robertson_mix1024.npz"))
    net = networks.robertson()
    model = sim.Model(net)
    nb = g["theta"].shape[0]
    out = _np(sim.simulate_batch(model, _solver(sim.SensitivityMethod.adjoint), g["theta"],
                                 np.zeros((nb, 0)), g["ts"], g["meas"],
                                 np.full(g["meas"].shape, float(g["sigma"]))))
    torch.cuda.synchronize()
    bad = np.nonzero(out["status"])[0]
    assert len(bad) == 0, f"conditions {bad.tolist()} failed: {out['status'][bad].tolist()}"
    ok = g["status"] == 0
    _gate(_normwise(out["sllh"][ok], g["sllh_asa"][ok]),
          _normwise(g["sllh_asa_rtol8"][ok], g["sllh_asa"][ok]), "sllh (adjoint, normwise)")
    nok = ~ok
    if nok.any():
        e = _normwise(out["sllh"][nok], g["sllh_asa_rtol8"][nok])
        assert (e <= 1e-3).all(), e
    cv = g["counters_rtol8"]
    st = out["stats"]
    for name, a, b in (("forward", st[:, 0], cv[:, 0]), ("backward", st[:, 8], cv[:, 9])):
        r = a.sum() / b.sum()
        print(f"{name} steps: GPU {a.mean():.1f} CVODES {b.mean():.1f} (ratio {r:.4f}); "
              f"instances with equal counts {(a == b).mean():.3f}")
        assert abs(r - 1.0) < 0.02, (name, r)


@pytest.mark.parametrize("name,fixture,meth", [("robertson", "robertson_batch64", 2),
                                               ("robertson", "robertson_batch64", 1),
                                               ("n40", "n40_conditions16", 2)])
def test_work_queue_backfill(gpu, name, fixture, meth):
    """With the grid capped below the batch size, every block takes several
    instances from the device work queue (re-initialising its LDS between
    them); the results equal those of the uncapped launch bit for bit."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / f"{fixture}.npz"))
    net = networks.robertson() if name == "robertson" else networks.network_n40()
    model = sim.Model(net)
    runs = []
    for cap in (0, 3):
        sol = _solver(meth, max_blocks=cap)
        out = _np(sim.simulate_batch(model, sol, g["theta"], g["kfix"], g["ts"], g["meas"],
                                     g["sigma"], want_sx=(meth == 1)))
        torch.cuda.synchronize()
        runs.append(out)
    assert (runs[0]["status"] == 0).all()
    for k in ("x", "y", "llh", "sllh", "status", "stats", "diag"):
        a, b = runs[0][k], runs[1][k]
        assert np.array_equal(a, b, equal_nan=True), k


def test_per_output_time_diagnostics(gpu):
    """numsteps / num_rhs_evals / ... per output time (ReturnData diagnostics,
    Solver::store_diagnosis / store_diagnosis_b): cumulative and monotone;
    the forward entries at the last output equal the run's counters; the
    backward entries start at zero (stored before the backward solver runs)
    and end at most at the backward run's total."""
    from amici_amd import sim
    g = dict(np.load(GOLD / "robertson_batch64.npz"))
    model = sim.Model(networks.robertson())
    sol = _solver(sim.SensitivityMethod.adjoint)
    edatas = []
    for b in range(4):
        e = sim.ExpData(model)
        e.set_timepoints(g["ts"])
        e.free_parameters = g["theta"][b]
        e.set_measurements(g["meas"][b])
        e.set_noise_scales(g["sigma"][b])
        edatas.append(e)
    for rd in sim.run_simulations(model, sol, edatas):
        assert rd.status == 0
        nt = len(rd.ts)
        for k in sim.DIAG_FIELDS:
            v = getattr(rd, k)
            assert v.shape == (nt,), k
            if k != "order":
                assert (np.diff(v) >= 0).all(), (k, v)
        assert rd.numsteps[-1] == rd.stats[0] and rd.num_rhs_evals[-1] == rd.stats[1]
        assert 1 <= rd.order[-1] <= 5
        assert rd.numsteps_b[0] == 0 and 0 < rd.numsteps_b[-1] <= rd.stats[8]


def test_stability_limit_detection(gpu):
    """BDF stability-limit detection (cvBDFStab / cvSLdet, AMICI's default)
    in the forward integration of all 1024 N40 conditions at the benchmark's
    tolerances, against CVODES's counters (fixture n40_stald1024): with the
    detector on, the total steps within 0.3 % and the order reductions it
    causes within 15 % of CVODES's; with it off, no reductions and the total
    steps within 0.3 % of CVODES's without the detector."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / "n40_stald1024.npz"))
    cs = networks.condition_set("N40")
    assert np.isclose(cs["theta"].sum(), g["theta_sum"], rtol=1e-12, atol=0)
    model = sim.Model(cs["network"])
    for flag, key in ((True, "on"), (False, "off")):
        sol = _solver()
        sol.set_stability_limit_flag(flag)
        out = _np(sim.simulate_batch(model, sol, cs["theta"], cs["kfix"], cs["ts"], want_x=False,
                                     want_y=False))
        torch.cuda.synchronize()
        assert (out["status"] == 0).all()
        cv = g[f"counters_{key}"]
        nst, nor = out["stats"][:, 0], out["stats"][:, 16]
        r = nst.sum() / cv[:, 0].sum()
        print(f"stldet {key}: steps GPU {nst.mean():.2f} CVODES {cv[:, 0].mean():.2f}; order "
              f"reductions GPU {nor.sum()} CVODES {cv[:, 15].sum()}; equal step counts "
              f"{(nst == cv[:, 0]).mean():.3f}")
        assert abs(r - 1.0) < 3e-3, (key, r)
        if flag:
            assert abs(nor.sum() - cv[:, 15].sum()) <= 0.15 * cv[:, 15].sum(), (nor.sum(), cv[:, 15].sum())
        else:
            assert nor.sum() == 0
