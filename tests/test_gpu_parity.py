"""GPU parity: the HIP batch path (through the C ABI) against the CVODES
oracle's golden fixtures (tests/golden/make_golden.py, rtol 1e-10).

Tolerances (stated per quantity, from the required accuracy of the path):
states and observables within 1e-6 relative, sllh within 1e-5 relative, each
with an absolute floor for entries that are zero in exact arithmetic (1e-9 of
the largest magnitude of x or y in that instance, 1e-7 of the largest sllh
entry); llh within the error that the observable tolerance implies
(sum of |dJy/dy| times the allowed observable error).  The GPU runs at the oracle's tolerances (rtol 1e-10,
atol 1e-14) so that differences are the two solvers' own truncation errors.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

from amici_amd import networks

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"
TIGHT = dict(rtol=1e-10, atol=1e-14, quad_rtol=1e-10, quad_atol=1e-14, maxsteps=200000,
             maxstepsB=400000, maxtime=60.0)

pytestmark = pytest.mark.gpu


def _load(name):
    return dict(np.load(GOLD / f"{name}.npz"))


def _close(a, b, rtol, floor_frac, what):
    a = np.asarray(a)
    b = np.asarray(b)
    floor = floor_frac * np.max(np.abs(b).reshape(b.shape[0], -1), axis=1)
    floor = floor.reshape((-1,) + (1,) * (b.ndim - 1))
    with np.errstate(invalid="ignore", divide="ignore"):
        err = np.where(a == b, 0.0, np.abs(a - b) / (rtol * np.abs(b) + floor))
    worst = np.unravel_index(np.argmax(err), err.shape)
    assert err.max() <= 1.0, (f"{what}: worst at {worst}: got {a[worst]!r} want {b[worst]!r} "
                              f"(ratio {err.max():.3g})")


Y_RTOL, Y_FLOOR = 1e-6, 1e-9


def _llh_tol(net, g):
    """Bound on |llh error| implied by the observable tolerance: the sum over
    measured points of |dJy/dy| times the allowed observable error."""
    y, m, s = g["y"], g["meas"], g["sigma"]
    sc = networks.noise_scale(net.noise)
    u = (lambda v: v) if sc == 0 else (np.log if sc == 1 else np.log10)
    dd = s if sc == 0 else (s * y if sc == 1 else s * y * np.log(10.0))
    r = (u(y) - u(m)) / s
    dj = (np.sign(r) if networks.noise_is_laplace(net.noise) else r) / dd
    dy = Y_RTOL * np.abs(y) + Y_FLOOR * np.abs(y).reshape(y.shape[0], -1).max(1)[:, None, None]
    return np.nansum(np.abs(dj) * dy, axis=(1, 2)) + 1e-12 * np.abs(g["llh"])


def _check_llh(net, out, g):
    tol = _llh_tol(net, g)
    err = np.abs(out["llh"] - g["llh"])
    assert (err <= tol).all(), f"llh: errors {err} above the implied bound {tol}"


def _dump(name, out):
    """With AMX_PARITY_DUMP=<dir>, keep the GPU results for offline analysis."""
    d = os.environ.get("AMX_PARITY_DUMP")
    if d:
        os.makedirs(d, exist_ok=True)
        np.savez_compressed(os.path.join(d, name + ".npz"),
                            **{k: v for k, v in out.items() if isinstance(v, np.ndarray)})


@pytest.fixture(scope="module")
def models(gpu):
    from amici_amd import sim
    return {"robertson": sim.Model(networks.robertson()),
            "n40": sim.Model(networks.network_n40()),
            "linear3": sim.Model(networks.linear_chain(3))}


def _run(models, model, g, meth, **over):
    import torch
    from amici_amd import sim
    sol = sim.Solver(**{**TIGHT, **over})
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    out = sim.simulate_batch(models[model], sol, g["theta"], g["kfix"], g["ts"], g["meas"],
                             g["sigma"], want_sx=(meth == sim.SensitivityMethod.forward))
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
            if k != "_keepalive"}


CASES = [("robertson", "robertson_nominal"), ("robertson", "robertson_batch64"),
         ("robertson", "robertson_edge"), ("linear3", "linear_chain3"),
         ("n40", "n40_conditions16")]


@pytest.mark.parametrize("model,fixture", CASES)
def test_states_observables(models, model, fixture):
    g = _load(fixture)
    out = _run(models, model, g, 0)
    _dump(f"states_{fixture}", out)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(out["y"], g["y"], Y_RTOL, Y_FLOOR, "y")
    _check_llh(models[model].network, out, g)


@pytest.mark.parametrize("model,fixture", CASES)
def test_forward_sensitivities(models, model, fixture):
    from amici_amd import sim
    g = _load(fixture)
    out = _run(models, model, g, sim.SensitivityMethod.forward)
    _dump(f"fsa_{fixture}", out)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _check_llh(models[model].network, out, g)
    _close(out["sllh"], g["sllh_fsa"], 1e-5, 1e-7, "sllh (forward)")
    if "sx" in g:
        _close(out["sx"], g["sx"], 1e-5, 1e-7, "sx")


@pytest.mark.parametrize("model,fixture", CASES)
def test_adjoint_gradient(models, model, fixture):
    from amici_amd import sim
    g = _load(fixture)
    out = _run(models, model, g, sim.SensitivityMethod.adjoint)
    _dump(f"asa_{fixture}", out)
    assert (out["status"] == 0).all(), out["status"]
    _check_llh(models[model].network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint)")
    # the oracle's own forward and adjoint gradients agree to the same bar
    _close(g["sllh_asa"], g["sllh_fsa"], 1e-5, 1e-7, "oracle sllh asa vs fsa")


def test_robertson_from_sbml(gpu):
    """Robertson written as SBML (amici_amd/data/robertson.xml), read by the
    SBML-subset reader, compiled and simulated, against the oracle's values."""
    from amici_amd import sbml, sim

    net = sbml.read_sbml(GOLD.parent.parent / "amici_amd" / "data" / "robertson.xml",
                         observables={"yA": "A", "yC": "C"})
    model = sim.Model(net)
    g = _load("robertson_nominal")
    out = _run({"m": model}, "m", g, sim.SensitivityMethod.forward)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(out["sllh"], g["sllh_fsa"], 1e-5, 1e-7, "sllh (forward)")


@pytest.fixture(scope="module")
def open_model(gpu):
    from amici_amd import sim
    return sim.Model(networks.network_n40_open())


@pytest.mark.parametrize("meth", [0, 1, 2])
def test_preequilibration(open_model, meth):
    """Pre-equilibration to steady state (damped Newton; implicit-function
    sensitivities) against the oracle's CVODES integration to t = 1e6 with
    forward sensitivities, for states, forward and adjoint gradients."""
    import torch
    from amici_amd import sim
    g = _load("n40open_preeq16")
    sol = sim.Solver(**TIGHT)
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    out = sim.simulate_batch(open_model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             kfix_preeq=g["kfix_preeq"], want_sx=(meth == 1))
    torch.cuda.synchronize()
    out = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
           if k != "_keepalive"}
    _dump(f"preeq_{meth}", out)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _check_llh(open_model.network, out, g)
    if meth == 1:
        _close(out["sx"][:2, 0], g["sx0"], 1e-5, 1e-7, "sx0 (implicit function)")
    if meth:
        key = "sllh_fsa" if meth == 1 else "sllh_asa"
        _close(out["sllh"], g[key], 1e-5, 1e-7, f"sllh ({meth})")


@pytest.fixture(scope="module")
def n200_model(gpu):
    from amici_amd import sim
    return sim.Model(networks.network_n200())


def test_n200_states(n200_model):
    """200-species network (one instance per 256-lane workgroup, blocked LU in
    global memory): states, observables and llh of 4 conditions."""
    g = _load("n200_conditions4")
    out = _run({"m": n200_model}, "m", g, 0)
    _dump("states_n200", out)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(out["y"], g["y"], Y_RTOL, Y_FLOOR, "y")
    _check_llh(n200_model.network, out, g)


def test_n200_adjoint(n200_model):
    from amici_amd import sim
    g = _load("n200_conditions4")
    out = _run({"m": n200_model}, "m", g, sim.SensitivityMethod.adjoint)
    _dump("asa_n200", out)
    assert (out["status"] == 0).all(), out["status"]
    _check_llh(n200_model.network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint)")


def test_n200_forward_sensitivities(n200_model):
    """Staggered forward sensitivities on the workgroup engine (each
    sensitivity's Newton correction solved with the global-memory LU) for a
    parameter list of 6 of the 600 parameters: sllh against the oracle's
    gradient (its adjoint, which agrees with its forward sensitivities to the
    same bar on the other fixtures), states against the oracle."""
    import torch
    from amici_amd import sim
    g = _load("n200_conditions4")
    plist = [0, 7, 150, 299, 451, 599]
    sol = sim.Solver(**TIGHT)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.forward)
    out = sim.simulate_batch(n200_model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             want_sx=True, plist=plist)
    torch.cuda.synchronize()
    out = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
           if k != "_keepalive"}
    _dump("fsa_n200", out)
    assert (out["status"] == 0).all(), out["status"]
    assert out["sllh"].shape == (4, len(plist)) and out["sx"].shape == (4, 40, len(plist), 200)
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _check_llh(n200_model.network, out, g)
    ref = g["sllh_asa"][:, plist]
    floor = 1e-7 * np.abs(g["sllh_asa"]).max(axis=1, keepdims=True)
    err = np.abs(out["sllh"] - ref) / (1e-5 * np.abs(ref) + floor)
    assert err.max() <= 1.0, (err.max(), out["sllh"], ref)


@pytest.mark.parametrize("meth", [0, 2])
def test_n200_preequilibration(gpu, meth):
    """Pre-equilibration of the open 200-species network on the workgroup
    engine (Newton steps and the adjoint's steady-state term through the
    global-memory LU of the Jacobian) against the oracle's CVODES integration
    to t = 1e6, for states and adjoint gradients."""
    import torch
    from amici_amd import sim
    model = sim.Model(networks.network_n200_open())
    g = _load("n200open_preeq4")
    sol = sim.Solver(**TIGHT)
    # AMICI's default acceptance (WRMS of dx/dt < 1 at 100 x the tolerances)
    # leaves a state error of about |dx/dt| / lambda_min, 1e-6 relative for
    # this network's slowest degradation rates (1e-2): the oracle's steady
    # state is integrated to t = 1e6 at rtol 1e-10, so the comparison asks
    # for a converged steady state
    sol.set_relative_tolerance_steady_state(1e-12)
    sol.set_absolute_tolerance_steady_state(1e-16)
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    out = sim.simulate_batch(model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             kfix_preeq=g["kfix_preeq"])
    torch.cuda.synchronize()
    out = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
           if k != "_keepalive"}
    _dump(f"preeq_n200_{meth}", out)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(out["y"], g["y"], Y_RTOL, Y_FLOOR, "y")
    _check_llh(model.network, out, g)
    if meth:
        _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint)")


@pytest.mark.parametrize("model,fixture", [("robertson", "robertson_batch64"),
                                           ("n40", "n40_conditions16")])
def test_adjoint_gradient_hermite(models, model, fixture):
    """The adjoint sweep with Hermite interpolation of the forward solution
    (InterpolationType hermite) against the oracle's adjoint gradient (AMICI's
    default polynomial interpolation): both interpolants are accurate to the
    integration tolerance, so the gradients agree to the parity bar."""
    from amici_amd import sim
    g = _load(fixture)
    out = _run(models, model, g, sim.SensitivityMethod.adjoint, interp=1)
    assert (out["status"] == 0).all(), out["status"]
    _check_llh(models[model].network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint, Hermite)")


def test_mix_composition(gpu):
    """The mixed workload's three model types at fixture size, run together
    on concurrent streams (sim.simulate_concurrent): 8 Robertson conditions of
    the mixed set (adjoint), 16 pre-equilibrated open 40-species conditions
    (adjoint) and 4 pre-equilibrated open 200-species conditions (adjoint),
    each group against its oracle fixture at the parity bar."""
    from amici_amd import sim
    gr, g40, g200 = _load("robertson_mix1024"), _load("n40open_preeq16"), _load("n200open_preeq4")
    idx = np.nonzero(gr["status"] == 0)[0][:8]
    nets = [networks.robertson(), networks.network_n40_open(), networks.network_n200_open()]

    def solver(steady=False):
        sol = sim.Solver(**TIGHT)
        if steady:
            sol.set_relative_tolerance_steady_state(1e-12)
            sol.set_absolute_tolerance_steady_state(1e-16)
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(sim.SensitivityMethod.adjoint)
        return sol

    nr = len(idx)
    groups = [
        dict(model=sim.Model(nets[0]), solver=solver(), theta=gr["theta"][idx], kfix=np.zeros((nr, 0)),
             ts=gr["ts"], meas=gr["meas"][idx], sigma=np.full(gr["meas"][idx].shape, float(gr["sigma"]))),
        dict(model=sim.Model(nets[1]), solver=solver(), theta=g40["theta"], kfix=g40["kfix"],
             ts=g40["ts"], meas=g40["meas"], sigma=g40["sigma"], kfix_preeq=g40["kfix_preeq"]),
        dict(model=sim.Model(nets[2]), solver=solver(steady=True), theta=g200["theta"],
             kfix=g200["kfix"], ts=g200["ts"], meas=g200["meas"], sigma=g200["sigma"],
             kfix_preeq=g200["kfix_preeq"]),
    ]
    outs = [{k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in o.items()
             if k != "_keepalive"} for o in sim.simulate_concurrent(groups)]
    for o in outs:
        assert (o["status"] == 0).all(), o["status"]
    r, o40, o200 = outs
    _close(r["sllh"], gr["sllh_asa"][idx], 1e-5, 1e-7, "Robertson sllh (adjoint)")
    np.testing.assert_allclose(r["llh"], gr["llh"][idx], rtol=1e-6)
    for o, g, net, what in ((o40, g40, nets[1], "n40open"), (o200, g200, nets[2], "n200open")):
        _close(o["x"], g["x"], Y_RTOL, Y_FLOOR, f"{what} x")
        _check_llh(net, o, g)
        _close(o["sllh"], g["sllh_asa"], 1e-5, 1e-7, f"{what} sllh (adjoint)")


@pytest.mark.parametrize("meth", [1, 2])
def test_parameter_list_scales_and_initial_state(models, meth):
    """Sensitivities with respect to a parameter list (5 parameters, not in
    order), parameters on mixed scales (none / ln / log10) and initial states
    given per condition (ExpData::plist / pscale / x0, applied by
    ConditionContext::apply_condition, src/edata.cpp:438-490): states, llh,
    sllh and sx against the oracle run with the same settings (fixture
    n40_plist_pscale4), through simulate_batch and through run_simulations."""
    from amici_amd import sim
    g = _load("n40_plist_pscale4")
    model = models["n40"]
    kw = dict(plist=g["plist"], pscale=g["pscale"], x0=g["x0"])
    out = _run_kw(model, g, meth, **kw)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _check_llh(model.network, out, g)
    assert out["sllh"].shape == (4, len(g["plist"]))
    key = "sllh_fsa" if meth == 1 else "sllh_asa"
    _close(out["sllh"], g[key], 1e-5, 1e-7, f"sllh ({meth})")
    if meth == 1:
        _close(out["sx"], g["sx"], 1e-5, 1e-7, "sx")
    # the same through ExpData and run_simulations
    sol = sim.Solver(**TIGHT)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(meth)
    eds = []
    for b in range(4):
        e = sim.ExpData(model)
        e.set_timepoints(g["ts"])
        e.fixed_parameters = g["kfix"][b]
        e.free_parameters = g["theta"][b]
        e.set_measurements(g["meas"][b])
        e.set_noise_scales(g["sigma"][b])
        e.plist, e.pscale, e.x0 = list(g["plist"]), list(g["pscale"]), g["x0"][b]
        eds.append(e)
    rds = sim.run_simulations(model, sol, eds)
    for b, rd in enumerate(rds):
        assert rd.status == 0
        np.testing.assert_array_equal(rd.sllh, out["sllh"][b])
        np.testing.assert_array_equal(rd.x, out["x"][b])


def _run_kw(model, g, meth, **kw):
    import torch
    from amici_amd import sim
    sol = sim.Solver(**TIGHT)
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    out = sim.simulate_batch(model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             want_sx=(meth == 1), **kw)
    torch.cuda.synchronize()
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
            if k != "_keepalive"}


@pytest.mark.parametrize("meth", [1, 2])
def test_initial_state_sensitivities(models, meth):
    """Caller-given initial state sensitivities (ExpData::sx0, with respect to
    the scaled parameters, for a parameter list [2, 0]) on 8 Robertson
    conditions: sllh and sx against the oracle (fixture robertson_sx0), and
    through Model.set_initial_state_sensitivities."""
    from amici_amd import sim
    g = _load("robertson_sx0")
    model = models["robertson"]
    out = _run_kw(model, g, meth, plist=g["plist"], sx0=g["sx0"])
    assert (out["status"] == 0).all(), out["status"]
    key = "sllh_fsa" if meth == 1 else "sllh_asa"
    _close(out["sllh"], g[key], 1e-5, 1e-7, f"sllh ({meth})")
    if meth == 1:
        _close(out["sx"], g["sx"], 1e-5, 1e-7, "sx")
    model.set_parameter_list(list(g["plist"]))
    try:
        model.set_initial_state_sensitivities(g["sx0"][0])
        one = _run_kw(model, {k: (v[:1] if k in ("theta", "kfix", "meas", "sigma") else v)
                              for k, v in g.items()}, meth)
        np.testing.assert_array_equal(one["sllh"][0], out["sllh"][0])
    finally:
        model.set_initial_state_sensitivities([])
        model.set_parameter_list(range(model.np))


@pytest.mark.parametrize("model,fixture,k,interp", [("robertson", "robertson_batch64", 32, 2),
                                                    ("robertson", "robertson_batch64", 5, 1),
                                                    ("n40", "n40_conditions16", 32, 2),
                                                    ("n40", "n40_conditions16", 7, 1)])
def test_adjoint_checkpoints(models, model, fixture, k, interp):
    """Adjoint with a checkpoint every k forward steps and each interval
    recomputed in the backward sweep (CVodeAdjInit(k), cvodea.c:99, with
    CVodeB's interval-wise sweep): llh and gradient against the oracle, which
    stores every step as AMICI does, at the parity bar; polynomial (2) and
    Hermite (1) interpolation."""
    from amici_amd import sim
    g = _load(fixture)
    out = _run(models, model, g, sim.SensitivityMethod.adjoint, ckpt_steps=k, interp=interp)
    _dump(f"asa_ck{k}_{fixture}", out)
    assert (out["status"] == 0).all(), out["status"]
    # every instance spans several checkpoint intervals
    assert (out["stats"][:, 0] > 3 * k).all(), out["stats"][:, 0]
    _check_llh(models[model].network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, f"sllh (adjoint, checkpoints every {k})")


@pytest.mark.parametrize("interp", [2, 1])
def test_adjoint_checkpoint_spacing_beyond_steps(models, interp):
    """A checkpoint spacing of max_traj (no checkpoint after the initial one,
    so nothing is recomputed) reproduces the run that stores every step: the
    forward pass bit for bit (llh, forward counters), the gradient to 1e-7
    relative (the backward kernels are compiled separately, and the compiler's
    contraction of multiply-adds may differ between them)."""
    from amici_amd import sim
    g = _load("n40_conditions16")
    base = _run(models, "n40", g, sim.SensitivityMethod.adjoint, interp=interp)
    ck = _run(models, "n40", g, sim.SensitivityMethod.adjoint, interp=interp,
              ckpt_steps=TIGHT["maxsteps"] + 1)
    assert (base["status"] == 0).all() and (ck["status"] == 0).all()
    assert np.array_equal(base["llh"], ck["llh"])
    assert np.array_equal(base["stats"][:, :8], ck["stats"][:, :8])
    _close(ck["sllh"], base["sllh"], 1e-7, 1e-9, "sllh (checkpoint kernels vs default)")


def test_n200_adjoint_checkpoints(n200_model):
    """The 200-species network with a checkpoint every 32 forward steps and
    recomputation (one instance per workgroup, the Newton matrix of the
    recomputation in the backward problem's global-memory slot) against the
    oracle's gradient."""
    from amici_amd import sim
    g = _load("n200_conditions4")
    out = _run({"m": n200_model}, "m", g, sim.SensitivityMethod.adjoint, ckpt_steps=32)
    _dump("asa_ck32_n200", out)
    assert (out["status"] == 0).all(), out["status"]
    assert (out["stats"][:, 0] > 3 * 32).all()
    _check_llh(n200_model.network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint, checkpoints every 32)")


def test_checkpoint_recomputation_replays_forward_steps(models):
    """The steps recomputed from the checkpoints in the backward sweep are the
    forward pass's own: for one traced instance, every accepted step and every
    failed attempt of each recomputed interval (event, time, step size, order,
    error norm) equals the forward pass's record bit for bit, interval by
    interval in the order of recomputation (the last interval first)."""
    import torch
    from amici_amd import sim
    K, inst, cap = 32, 3, 1 << 16
    g = _load("n40_conditions16")
    lib = models["n40"].lib.lib
    buf = torch.zeros((cap, 8), dtype=torch.float64, device="cuda")
    assert lib.amx_set_trace(buf.data_ptr(), inst, cap) == 0
    try:
        out = _run(models, "n40", g, sim.SensitivityMethod.adjoint, ckpt_steps=K)
    finally:
        lib.amx_set_trace(None, 0, 0)
    assert (out["status"] == 0).all(), out["status"]
    tr = buf.cpu().numpy()
    fwd = tr[: cap // 2]
    rec = tr[cap // 2 + cap // 4:]
    fwd, rec = fwd[fwd[:, 0] > 0], rec[rec[:, 0] > 0]
    nst = int(out["stats"][inst, 0])
    assert int((fwd[:, 0] == 1).sum()) == nst
    # forward records grouped by checkpoint interval: interval c holds the
    # accepted steps c K + 1 .. (c + 1) K and the failed attempts before them
    step_no = np.cumsum(fwd[:, 0] == 1) + (fwd[:, 0] != 1)
    interval = (step_no - 1) // K
    nrec = nst // K  # intervals before the last one, which the forward pass kept
    assert nrec >= 3
    want = np.concatenate([fwd[interval == c] for c in range(nrec - 1, -1, -1)])
    assert rec.shape == want.shape, (rec.shape, want.shape)
    assert np.array_equal(rec[:, :5], want[:, :5])


def test_checkpoints_with_preequilibration(open_model):
    """Checkpoints every 16 steps after a pre-equilibration: the recomputed
    intervals start from the steady state (record 0), and the adjoint's
    steady-state term is added at t0 as without checkpoints; gradient against
    the oracle."""
    import torch
    from amici_amd import sim
    g = _load("n40open_preeq16")
    sol = sim.Solver(**TIGHT, ckpt_steps=16)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    out = sim.simulate_batch(open_model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             kfix_preeq=g["kfix_preeq"])
    torch.cuda.synchronize()
    out = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
           if k != "_keepalive"}
    assert (out["status"] == 0).all(), out["status"]
    assert (out["stats"][:, 0] > 3 * 16).all()
    _check_llh(open_model.network, out, g)
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (adjoint, pre-equilibrated, checkpoints)")


def test_checkpoint_step_limit(models):
    """With checkpoints max_traj still bounds the forward points of an
    instance: a limit below the steps taken fails every instance with the
    trajectory-buffer status (-1010) and NaN gradients, and a limit above
    them changes nothing."""
    from amici_amd import sim
    g = _load("n40_conditions16")
    small = _run(models, "n40", g, sim.SensitivityMethod.adjoint, ckpt_steps=32, max_traj=200)
    assert (small["status"] == -1010).all(), small["status"]
    assert np.isnan(small["sllh"]).all()
    big = _run(models, "n40", g, sim.SensitivityMethod.adjoint, ckpt_steps=32, max_traj=50000)
    ref = _run(models, "n40", g, sim.SensitivityMethod.adjoint, ckpt_steps=32)
    assert (ref["status"] == 0).all()
    assert np.array_equal(big["sllh"], ref["sllh"]) and np.array_equal(big["llh"], ref["llh"])


def test_n200_preequilibration_forward_sensitivities(gpu):
    """Forward sensitivities after a pre-equilibration on the workgroup
    engine: the steady state's implicit-function sensitivities come from the
    global-memory LU of the Jacobian; a parameter list of 5 parameters against
    the oracle's gradient."""
    import torch
    from amici_amd import sim
    model = sim.Model(networks.network_n200_open())
    g = _load("n200open_preeq4")
    plist = [0, 3, 200, 512, model.np - 1]
    sol = sim.Solver(**TIGHT)
    sol.set_relative_tolerance_steady_state(1e-12)
    sol.set_absolute_tolerance_steady_state(1e-16)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.forward)
    out = sim.simulate_batch(model, sol, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"],
                             kfix_preeq=g["kfix_preeq"], plist=plist)
    torch.cuda.synchronize()
    out = {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()
           if k != "_keepalive"}
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    ref = g["sllh_asa"][:, plist]
    floor = 1e-7 * np.abs(g["sllh_asa"]).max(axis=1, keepdims=True)
    err = np.abs(out["sllh"] - ref) / (1e-5 * np.abs(ref) + floor)
    assert err.max() <= 1.0, (err.max(), out["sllh"], ref)


@pytest.mark.parametrize("noise", [2, 3, 4, 5])
def test_noise_models(gpu, noise):
    """Robertson with the log10-normal, Laplace, log-Laplace and log10-Laplace
    noise models (each its own model module): states, llh, forward and adjoint
    gradients against the oracle."""
    from amici_amd import sim
    g = _load(f"robertson_noise{noise}")
    net = networks.with_noise(networks.robertson(), noise)
    models = {"m": sim.Model(net)}
    for meth, key in ((sim.SensitivityMethod.forward, "sllh_fsa"),
                      (sim.SensitivityMethod.adjoint, "sllh_asa")):
        out = _run(models, "m", g, meth)
        assert (out["status"] == 0).all(), out["status"]
        _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
        _check_llh(net, out, g)
        _close(out["sllh"], g[key], 1e-5, 1e-7, f"sllh ({key}, noise {noise})")


@pytest.mark.parametrize("model,fixture", [("robertson", "robertson_batch64"),
                                           ("n40", "n40_conditions16")])
def test_start_time(models, model, fixture):
    """Start times other than 0 (ExpData t_start, Model.set_t0): the models
    are autonomous, so a run started at t0 with every output time shifted by
    t0 must give the unshifted run's states, llh and gradients; checked
    against the oracle's fixture at the parity bar with a different t0 per
    instance, for states, forward and adjoint sensitivities."""
    from amici_amd import sim
    g = _load(fixture)
    nb = g["theta"].shape[0]
    t0 = np.linspace(0.5, 30.0, nb)
    shifted = dict(g, ts=g["ts"][None, :] + t0[:, None])
    for meth in (0, sim.SensitivityMethod.forward, sim.SensitivityMethod.adjoint):
        out = _run_kw(models[model], shifted, meth, t0=t0)
        assert (out["status"] == 0).all(), out["status"]
        _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
        _close(out["y"], g["y"], Y_RTOL, Y_FLOOR, "y")
        _check_llh(models[model].network, out, g)
        if meth:
            key = "sllh_fsa" if meth == sim.SensitivityMethod.forward else "sllh_asa"
            _close(out["sllh"], g[key], 1e-5, 1e-7, f"sllh ({key})")
    # the plain one-instance-per-launch path
    import torch
    seq = sim.simulate_batch(models[model], sim.Solver(**TIGHT), g["theta"][:4], g["kfix"][:4],
                             shifted["ts"][:4], sequential=True, t0=t0[:4])
    torch.cuda.synchronize()
    _close(seq["x"].cpu().numpy(), g["x"][:4], Y_RTOL, Y_FLOOR, "x (sequential)")
    # the same through ExpData.t_start and run_simulations
    m = models[model]
    sol = sim.Solver(**TIGHT)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    eds = []
    for i in range(4):
        e = sim.ExpData(m)
        e.set_timepoints(shifted["ts"][i])
        e.t_start = float(t0[i])
        e.free_parameters = g["theta"][i]
        e.fixed_parameters = g["kfix"][i]
        e.set_measurements(g["meas"][i])
        e.set_noise_scales(g["sigma"][i])
        eds.append(e)
    rd = sim.run_simulations(m, sol, eds)
    for i, r in enumerate(rd):
        assert r.status == 0
        _close(np.asarray(r.sllh)[None], g["sllh_asa"][i:i + 1], 1e-5, 1e-7, "sllh (run_simulations)")


def test_presimulation(models):
    """Pre-simulation (ExpData t_presim, fixed_parameters_presimulation):
    the condition runs from t0 - t_presim with the pre-simulation's fixed
    parameters, which set the input species' initial amounts, and continues
    through t0 without reinitialisation.  With t_presim two output intervals
    of the N40 grid, the states at ts[k] are the oracle's states at ts[k + 2]
    of the run with the pre-simulation inputs; the condition's own fixed
    parameters (here deliberately different) play no part."""
    from amici_amd import sim
    g = _load("n40_conditions16")
    m = models["n40"]
    ts = g["ts"]
    T = ts[2] - ts[0]
    sol = sim.Solver(**TIGHT)
    eds = []
    for i in range(8):
        e = sim.ExpData(m)
        e.set_timepoints(ts[:-2])
        e.free_parameters = g["theta"][i]
        e.fixed_parameters = g["kfix"][(i + 1) % 8] * 3.0
        e.fixed_parameters_presimulation = g["kfix"][i]
        e.t_presim = T
        eds.append(e)
    rd = sim.run_simulations(m, sol, eds)
    x = np.stack([r.x for r in rd])
    assert all(r.status == 0 for r in rd)
    _close(x, g["x"][:8, 2:], Y_RTOL, Y_FLOOR, "x after pre-simulation")


def test_model_start_time_default(models):
    """Model.set_t0 applies to every condition of a batch that gives no start
    time of its own (ExpData t_start overrides it, src/edata.cpp:551)."""
    from amici_amd import sim
    g = _load("robertson_batch64")
    m = sim.Model(networks.robertson())
    m.set_t0(12.5)
    shifted = dict(g, ts=g["ts"] + 12.5)
    out = _run_kw(m, shifted, sim.SensitivityMethod.adjoint)
    assert (out["status"] == 0).all(), out["status"]
    _close(out["x"], g["x"], Y_RTOL, Y_FLOOR, "x")
    _close(out["sllh"], g["sllh_asa"], 1e-5, 1e-7, "sllh (model t0)")
    # a condition's own start time wins
    e = sim.ExpData(m)
    e.set_timepoints(g["ts"])
    e.t_start = 0.0
    e.free_parameters = g["theta"][0]
    e.set_measurements(g["meas"][0])
    e.set_noise_scales(g["sigma"][0])
    sol = sim.Solver(**TIGHT)
    r = sim.run_simulations(m, sol, [e])[0]
    assert r.status == 0
    _close(np.asarray(r.x)[None], g["x"][:1], Y_RTOL, Y_FLOOR, "x (ExpData t_start 0)")
