"""The drop-in entry point ``run_simulations(model, solver, edatas)`` on the
GPU: ExpData in, ReturnData out, as a user of the reference calls it.

Covers the ReturnData fields derived on top of the kernels' outputs (res,
sres, chi2, FIM, sy; src/rdata.cpp:985-1204) through identities the
likelihood fixes (Gauss-Newton gradient, error residuals completing the
log-likelihood), the reporting modes (src/rdata.cpp:85-160), failfast
(src/amici.cpp:246-260), the step budget of Solver::apply_max_num_steps
(src/solver.cpp:107-119, fixture tests/golden/robertson_step_limit.npz from
the oracle) and the reference's own test that residuals of a failed
simulation are invalidated (python/tests/test_rdata.py:30-64).
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from amici_amd import networks
from test_gpu_parity import TIGHT, Y_FLOOR, Y_RTOL, _close, _llh_tol, _load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sim_models(gpu):
    from amici_amd import sim
    return {"robertson": sim.Model(networks.robertson()),
            "n40": sim.Model(networks.network_n40())}


def _solver(meth=0, **over):
    from amici_amd import sim
    s = sim.Solver(**{**TIGHT, **over})
    if meth:
        s.set_sensitivity_order(sim.SensitivityOrder.first)
        s.set_sensitivity_method(meth)
    return s


def _edatas(model, g, idx=None):
    from amici_amd import sim
    idx = range(g["theta"].shape[0]) if idx is None else idx
    out = []
    for b in idx:
        e = sim.ExpData(model)
        e.set_timepoints(g["ts"])
        e.fixed_parameters = g["kfix"][b]
        e.free_parameters = g["theta"][b]
        e.set_measurements(g["meas"][b])
        e.set_noise_scales(g["sigma"][b])
        out.append(e)
    return out


def _stack(rdatas, key):
    return np.stack([getattr(r, key) for r in rdatas])


def test_forward_sensitivities_and_derived_fields(sim_models):
    from amici_amd import sim
    model = sim_models["n40"]
    g = _load("n40_conditions16")
    rds = sim.run_simulations(model, _solver(sim.SensitivityMethod.forward), _edatas(model, g))
    assert [r.status for r in rds] == [0] * len(rds)
    _close(_stack(rds, "x"), g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(_stack(rds, "y"), g["y"], Y_RTOL, Y_FLOOR, "y")
    llh = _stack(rds, "llh")
    assert (np.abs(llh - g["llh"]) <= _llh_tol(model.network, g)).all()
    sllh = _stack(rds, "sllh")
    _close(sllh, g["sllh_fsa"], 1e-5, 1e-7, "sllh")
    # sy = C sx, [nt, np, ny]
    C = model.network.observable_matrix
    sy_host = np.einsum("btpx,yx->btpy", _stack(rds, "sx"), C)
    np.testing.assert_allclose(_stack(rds, "sy"), sy_host, rtol=1e-12, atol=1e-14 * np.abs(sy_host).max())
    # sigma from the data: sllh = -sum res * sres (Gauss-Newton gradient, exact
    # for normal and log-normal noise), the kernel's own accumulation against
    # the host-side fields
    res, sres = _stack(rds, "res"), _stack(rds, "sres")
    assert res.shape == (16, 25 * 30) and sres.shape == (16, 25 * 30, model.np)
    _close(-np.einsum("bk,bkp->bp", res, sres), sllh, 1e-9, 1e-12, "Gauss-Newton gradient")
    fim = _stack(rds, "FIM")
    np.testing.assert_allclose(fim, np.einsum("bki,bkj->bij", sres, sres), rtol=1e-12,
                               atol=1e-12 * np.abs(fim).max())
    assert np.all(np.linalg.eigvalsh(fim) > -1e-9 * np.abs(fim).max())
    # error residuals complete the log-likelihood: with
    # res_err = sqrt(2 log sigma + c), -llh = chi2 / 2 + n log(2 pi) / 2, plus
    # the sum of log m for log-normal noise (the density's 1/m factor)
    model.set_add_sigma_residuals(True)
    try:
        rs = sim.run_simulations(model, _solver(sim.SensitivityMethod.forward), _edatas(model, g))
    finally:
        model.set_add_sigma_residuals(False)
    n = (~np.isnan(g["meas"])).reshape(16, -1).sum(1)
    const = 0.5 * n * math.log(2 * math.pi)
    if model.network.noise == networks.NOISE_LOGNORMAL:
        const = const + np.nansum(np.log(g["meas"]), axis=(1, 2))
    chi2 = _stack(rs, "chi2")
    np.testing.assert_allclose(-_stack(rs, "llh"), 0.5 * chi2 + const, rtol=1e-12)
    assert _stack(rs, "res").shape == (16, 2 * 25 * 30)
    np.testing.assert_array_equal(_stack(rs, "res")[:, :750], res)


def test_adjoint_through_run_simulations(sim_models):
    from amici_amd import sim
    model = sim_models["robertson"]
    g = _load("robertson_batch64")
    rds = sim.run_simulations(model, _solver(sim.SensitivityMethod.adjoint), _edatas(model, g))
    assert [r.status for r in rds] == [0] * len(rds)
    _close(_stack(rds, "sllh"), g["sllh_asa"], 1e-5, 1e-7, "sllh")
    assert all(r.sx is None and r.sy is None and r.sres is None and r.FIM is None for r in rds)
    # -llh = chi2 / 2 + sum over measured of log(2 pi sigma^2) / 2, with
    # log m added per point for log-normal noise
    m, s = g["meas"], g["sigma"]
    const = 0.5 * np.nansum(np.log(2 * math.pi * s * s) + 0 * m, axis=(1, 2))
    if model.network.noise == networks.NOISE_LOGNORMAL:
        const = const + np.nansum(np.log(m), axis=(1, 2))
    np.testing.assert_allclose(-_stack(rds, "llh"), 0.5 * _stack(rds, "chi2") + const, rtol=1e-11)


def test_reporting_modes(sim_models):
    from amici_amd import sim
    model = sim_models["n40"]
    g = _load("n40_conditions16")
    ed = _edatas(model, g, range(4))
    out = {}
    for mode in sim.RDataReporting:
        s = _solver(sim.SensitivityMethod.forward)
        s.set_return_data_reporting_mode(mode)
        out[mode] = sim.run_simulations(model, s, ed)
    full = out[sim.RDataReporting.full]
    lik = out[sim.RDataReporting.likelihood]
    res = out[sim.RDataReporting.residuals]
    ol = out[sim.RDataReporting.observables_likelihood]
    for b in range(4):
        assert lik[b].x is None and lik[b].y is None and lik[b].res is None
        assert lik[b].llh == full[b].llh
        np.testing.assert_array_equal(lik[b].FIM, full[b].FIM)
        assert math.isnan(res[b].llh) and res[b].sllh is None and res[b].x is None
        np.testing.assert_array_equal(res[b].res, full[b].res)
        np.testing.assert_array_equal(res[b].sres, full[b].sres)
        assert ol[b].x is None and ol[b].res is None
        np.testing.assert_array_equal(ol[b].y, full[b].y)
        np.testing.assert_array_equal(ol[b].sy, full[b].sy)


def test_invalidate_residuals_on_failure(sim_models):
    """python/tests/test_rdata.py:30-64 on Robertson (normal noise, so the
    synthetic data y + N(0, 1) may be negative)."""
    from amici_amd import sim
    model = sim_models["robertson"]
    model.set_timepoints(np.linspace(0, 60, 61))
    solver = model.create_solver()
    solver.set_sensitivity_method(sim.SensitivityMethod.forward)
    solver.set_sensitivity_order(sim.SensitivityOrder.first)
    edata = sim.ExpData(sim.run_simulation(model, solver), 1.0, 0.0, seed=1)
    rdata = sim.run_simulation(model, solver, edata)
    assert rdata.status == sim.AMICI_SUCCESS
    assert np.all(np.isfinite(rdata.res))
    assert np.all(np.isfinite(rdata.sres))
    assert np.all(np.isfinite(rdata.FIM))
    solver.set_max_steps(1)
    rdata = sim.run_simulation(model, solver, edata)
    assert rdata.status != sim.AMICI_SUCCESS
    assert np.isnan(rdata.llh)
    nytrue = len(model.get_observable_ids())
    assert np.all(np.isfinite(rdata.res[:nytrue]))
    assert np.all(np.isnan(rdata.res[nytrue:]))
    assert np.all(np.isnan(rdata.sres[nytrue:, :]))
    assert np.all(np.isnan(rdata.FIM))


def _first_nan_row(x):
    bad = np.isnan(x).any(-1)
    return int(np.argmax(bad)) if bad.any() else len(x)


def test_step_budget(sim_models):
    """Statuses and the first output time not reached for forward step
    budgets: exhausted inside an interval -> AMICI_TOO_MUCH_WORK, exactly at
    an output time -> AMICI_ERROR.  The GPU's own cumulative step counts per
    output time (one batch over truncated grids, padded by repeating the last
    time) decide which case a budget is; where they equal the oracle's, status
    and output index must equal the oracle's too."""
    import torch
    from amici_amd import sim
    model = sim_models["robertson"]
    g = _load("robertson_step_limit")
    ts, nt = g["ts"], len(g["ts"])
    grids = np.stack([np.concatenate([ts[:k], np.full(nt - k, ts[k - 1])]) for k in range(1, nt + 1)])
    out = sim.simulate_batch(model, _solver(), np.repeat(g["theta"], nt, 0), np.zeros((nt, 0)), grids)
    torch.cuda.synchronize()
    assert (out["status"].cpu().numpy() == 0).all()
    steps = out["stats"].cpu().numpy()[:, 0].astype(np.int64)
    assert np.all(np.diff(steps) > 0)
    budgets = sorted(set(int(m) for m in g["budgets"]) | {int(steps[5]), int(steps[5]) + 1})
    for m in budgets:
        e = sim.ExpData(model)
        e.set_timepoints(ts)
        rd = sim.run_simulation(model, _solver(maxsteps=m), e)
        want = -99 if m in steps else -1
        assert rd.status == want, (m, rd.status)
        assert _first_nan_row(rd.x) == int((steps <= m).sum()), m
        assert math.isnan(rd.llh)
        assert int(rd.stats[0]) == m
        j = np.nonzero(g["budgets"] == m)[0]
        if len(j) and np.array_equal(steps[steps <= m + 1], g["steps"][g["steps"] <= m + 1]):
            assert rd.status == int(g["status"][j[0]])
            assert _first_nan_row(rd.x) == int(g["first_bad"][j[0]])


def test_failfast_and_mixed_grids(sim_models):
    from amici_amd import sim
    model = sim_models["robertson"]
    short = np.logspace(-4.0, 0.0, 40)
    long = networks.robertson_times()
    ed = []
    for ts in (short, long, short):
        e = sim.ExpData(model)
        e.set_timepoints(ts)
        e.set_measurements(np.full((40, model.ny), 0.5))
        e.set_noise_scales(0.1)
        ed.append(e)
    solver = _solver(maxsteps=500)
    r_ff = sim.run_simulations(model, solver, ed, failfast=True)
    r_all = sim.run_simulations(model, solver, ed, failfast=False)
    assert [r.status for r in r_ff] == [0, -1, sim.AMICI_NOT_RUN]
    assert [r.status for r in r_all] == [0, -1, 0]
    bad = _first_nan_row(r_all[1].x)
    assert 0 < bad < 40
    res = r_all[1].res.reshape(40, model.ny)
    assert np.isfinite(res[:bad]).all() and np.isnan(res[bad:]).all()
    assert math.isnan(r_all[1].chi2) and math.isfinite(r_all[0].chi2)
    np.testing.assert_array_equal(r_all[0].x, r_all[2].x)
    # conditions on grids of different lengths are batched per length and
    # returned in order, with the same results as separate calls
    e_short = sim.ExpData(model)
    e_short.set_timepoints(short[:10])
    mixed = sim.run_simulations(model, _solver(), [ed[0], e_short, ed[2]])
    alone = sim.run_simulation(model, _solver(), e_short)
    np.testing.assert_array_equal(mixed[1].x, alone.x)
    np.testing.assert_array_equal(mixed[0].x, r_all[0].x)


def test_concurrent_groups_equal_sequential(sim_models):
    """Two models' batches on concurrent streams give the same results, bit
    for bit, as the batches run one after another."""
    import torch
    from amici_amd import sim
    gr, gn = _load("robertson_batch64"), _load("n40_conditions16")
    sol = _solver(sim.SensitivityMethod.adjoint, rtol=1e-8, atol=1e-12, quad_rtol=1e-8,
                  quad_atol=1e-12)
    groups = [dict(model=sim_models["robertson"], solver=sol, theta=gr["theta"], kfix=gr["kfix"],
                   ts=gr["ts"], meas=gr["meas"], sigma=gr["sigma"]),
              dict(model=sim_models["n40"], solver=sol, theta=gn["theta"], kfix=gn["kfix"],
                   ts=gn["ts"], meas=gn["meas"], sigma=gn["sigma"])]
    conc = sim.simulate_concurrent(groups)
    for g, c in zip(groups, conc):
        kw = dict(g)
        alone = sim.simulate_batch(kw.pop("model"), kw.pop("solver"), kw.pop("theta"), kw.pop("kfix"),
                                   kw.pop("ts"), **kw)
        torch.cuda.synchronize()
        for k in ("llh", "sllh", "status", "x"):
            np.testing.assert_array_equal(c[k].cpu().numpy(), alone[k].cpu().numpy(), err_msg=k)
    # one model on two concurrent streams: each stream has its own workspace
    # (work-queue counters, trajectory), so both batches equal the sequential one
    twice = sim.simulate_concurrent([groups[0], groups[0]])
    for c in twice:
        for k in ("llh", "sllh", "status", "x"):
            np.testing.assert_array_equal(c[k].cpu().numpy(), conc[0][k].cpu().numpy(), err_msg=k)


def test_model_simulate(sim_models):
    """Model.simulate (python/sdist/amici/sim/sundials/_swig_wrappers.py:296-331):
    a list of ExpData runs run_simulations, one ExpData run_simulation; the
    sensitivity settings apply to the call only."""
    from amici_amd import sim
    model = sim_models["robertson"]
    g = _load("robertson_batch64")
    ed = _edatas(model, g, range(4))
    solver = _solver()
    many = model.simulate(solver=solver, edata=ed, sensi_method="adjoint", sensi_order="first")
    ref = sim.run_simulations(model, _solver(sim.SensitivityMethod.adjoint), ed)
    for a, b in zip(many, ref):
        np.testing.assert_array_equal(a.sllh, b.sllh)
    assert solver.get_sensitivity_order() == sim.SensitivityOrder.none
    one = model.simulate(solver=solver, edata=ed[1])
    assert one.sllh is None and one.status == 0
    np.testing.assert_array_equal(one.x, ref[1].x)


@pytest.mark.parametrize("meth", [0, 2])
def test_failfast_skips_on_device(sim_models, meth):
    """With failfast the kernels skip the instances with a larger index than a
    failed one that have not started yet (run_simulations' skipThrough,
    src/amici.cpp:232-260).  One block of four teams works through 24
    Robertson conditions; condition 0 fails at once (NaN parameters), so the
    conditions its team and the others take from the queue afterwards never
    run (no steps, status AMICI_NOT_RUN); the conditions that ran equal a run
    without failfast, and run_simulations reports every condition after the
    first failure as not run."""
    import torch
    from amici_amd import sim
    model = sim_models["robertson"]
    g = _load("robertson_batch64")
    nb = 24
    theta = g["theta"][:nb].copy()
    theta[0] = np.nan
    args = (theta, g["kfix"][:nb], g["ts"], g["meas"][:nb], g["sigma"][:nb])
    sol = _solver(meth, max_blocks=1, rtol=1e-8, atol=1e-12)
    ff = sim.simulate_batch(model, sol, *args, failfast=True)
    ref = sim.simulate_batch(model, sol, *args)
    torch.cuda.synchronize()
    st, stats = ff["status"].cpu().numpy(), ff["stats"].cpu().numpy()
    assert st[0] < 0 and ref["status"].cpu().numpy()[0] == st[0]
    skipped = np.nonzero(st == sim.AMICI_NOT_RUN)[0]
    assert len(skipped) >= nb - 8 and skipped.min() >= 1, st
    assert (stats[skipped, 0] == 0).all() and np.isnan(ff["llh"].cpu().numpy()[skipped]).all()
    ran = np.nonzero(st == 0)[0]
    np.testing.assert_array_equal(ff["llh"].cpu().numpy()[ran], ref["llh"].cpu().numpy()[ran])
    if meth:
        np.testing.assert_array_equal(ff["sllh"].cpu().numpy()[ran], ref["sllh"].cpu().numpy()[ran])
    # run_simulations: serial semantics on top
    eds = []
    for b in range(nb):
        e = sim.ExpData(model)
        e.set_timepoints(g["ts"])
        e.free_parameters = theta[b]
        e.set_measurements(g["meas"][b])
        e.set_noise_scales(g["sigma"][b])
        eds.append(e)
    rds = sim.run_simulations(model, sol, eds, failfast=True)
    assert rds[0].status < 0 and all(r.status == sim.AMICI_NOT_RUN for r in rds[1:])
