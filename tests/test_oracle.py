"""Pins the oracle (reference CVODES + restated AMICI layer) before it is
trusted as the checker.

Each restated piece is checked against something the mathematics fixes:
closed forms (linear ODE and its parameter derivatives via the block matrix
exponential), invariants (mass conservation, replicate identity), published
reference values (Robertson), library routines (scipy Radau, numpy), sympy
derivatives and central finite differences of the log-likelihood (the recipe of
python/sdist/amici/sim/sundials/gradient_check.py:106-129, epsilon on the log10
scale).
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp
from scipy.linalg import expm

from amici_amd import codegen, networks
from oracle import cvodes_oracle as O

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"


def _linear_A(net, theta):
    """xdot = A x for a first-order network."""
    k = 10.0 ** np.asarray(theta)
    A = np.zeros((net.nx, net.nx))
    for rx in net.reactions:
        (a,), (b,) = rx.reactants, rx.products
        A[a, a] -= k[rx.rate]
        A[b, a] += k[rx.rate]
    return A, k


def test_xdot_jac_dfdp_match_sympy_and_fd():
    rng = np.random.default_rng(0)
    for net in (networks.robertson(), networks.network_n40(), networks.linear_chain(4)):
        om = O.OracleModel(net)
        sm = codegen.symbolic_model(net)
        theta = net.theta_nominal
        p = 10.0 ** theta
        x = rng.uniform(0.1, 2.0, net.nx)
        subs = {**dict(zip(sm["x"], x)), **dict(zip(sm["p"], p))}
        f_sym = np.array([float(e.subs(subs)) for e in sm["xdot"]])
        np.testing.assert_allclose(om.xdot(theta, x), f_sym, rtol=1e-12, atol=1e-12 * np.abs(f_sym).max())
        J = om.jac(theta, x)
        J_sym = np.array(sm["J"].subs(subs).evalf(), dtype=float)
        np.testing.assert_allclose(J, J_sym, rtol=1e-12, atol=1e-12 * np.abs(J_sym).max())
        # dxdot/dp against sympy and against a finite difference in p
        dfdp = om.dxdotdp(theta, x)
        dfdp_sym = np.array(sm["dfdp"].subs(subs).evalf(), dtype=float)
        np.testing.assert_allclose(dfdp, dfdp_sym, rtol=1e-12, atol=1e-12 * np.abs(dfdp_sym).max())
        # xdot is linear in each rate constant, so a step of p_j itself is an
        # exact difference quotient up to roundoff
        scale = np.abs(om.xdot(theta, x)).max() + np.abs(J @ x).max()
        for j in range(net.np):
            tp = theta.copy()
            tp[j] = np.log10(2.0 * p[j])
            fd = (om.xdot(tp, x) - om.xdot(theta, x)) / p[j]
            np.testing.assert_allclose(dfdp[:, j], fd, rtol=1e-9, atol=1e-12 * scale / p[j])


def test_noise_closed_forms():
    s, m = 0.1, 1.7
    for y in (0.3, 1.7, 2.9):
        jn = O.Jy(0, y, s, m)
        assert jn == pytest.approx(0.5 * math.log(2 * math.pi * s * s) + 0.5 * ((y - m) / s) ** 2, rel=1e-14)
        jl = O.Jy(1, y, s, m)
        assert jl == pytest.approx(0.5 * math.log(2 * math.pi * s * s * m * m)
                                   + 0.5 * ((math.log(y) - math.log(m)) / s) ** 2, rel=1e-14)
        for noise in (0, 1):
            h = 1e-6
            fd = (O.Jy(noise, y + h, s, m) - O.Jy(noise, y - h, s, m)) / (2 * h)
            assert O.dJydy(noise, y, s, m) == pytest.approx(fd, rel=1e-7, abs=1e-9)


@pytest.mark.parametrize("noise", [0, 1, 2, 3, 4, 5])
def test_noise_models_are_log_densities(noise):
    """Jy of every continuous noise model is minus the log density of the
    measurement m given y and sigma (scipy.stats: Gaussian or Laplace on the
    linear, ln or log10 scale, with the scale's Jacobian 1 / (m ln b)), it
    integrates to one over m, and dJy/dy is its derivative."""
    from scipy import integrate, stats
    s = 0.3
    sc = networks.noise_scale(noise)
    lap = networks.noise_is_laplace(noise)
    u = (lambda v: v) if sc == 0 else (np.log if sc == 1 else np.log10)
    jac = (lambda m: 1.0) if sc == 0 else ((lambda m: m) if sc == 1 else (lambda m: m * math.log(10)))
    dist = stats.laplace if lap else stats.norm
    for y in (0.3, 1.7, 2.9):
        for m in (0.2, 1.7, 3.5):
            want = -(dist.logpdf(u(m), loc=u(y), scale=s) - math.log(jac(m)))
            assert O.Jy(noise, y, s, m) == pytest.approx(want, rel=1e-12, abs=1e-12)
            if m != y:
                h = 1e-6
                fd = (O.Jy(noise, y + h, s, m) - O.Jy(noise, y - h, s, m)) / (2 * h)
                assert O.dJydy(noise, y, s, m) == pytest.approx(fd, rel=1e-6, abs=1e-8)
        lo = -np.inf if sc == 0 else 0.0
        total = sum(integrate.quad(lambda m: math.exp(-O.Jy(noise, y, s, m)), a, b, limit=200)[0]
                    for a, b in ((lo, y), (y, np.inf)))
        assert total == pytest.approx(1.0, rel=1e-7)


def test_linear_states_and_fsa_closed_form():
    net = networks.linear_chain(3)
    ts = np.array([0.0, 0.1, 0.5, 1.0, 3.0, 10.0])
    r = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16,
                   sensi=1, want_sx=True)
    assert r["status"][0] == 0
    A, k = _linear_A(net, net.theta_nominal)
    x0 = net.x0_const
    n = net.nx
    for it, t in enumerate(ts):
        np.testing.assert_allclose(r["x"][0, it], expm(A * t) @ x0, rtol=1e-9, atol=1e-13)
        for j in range(net.np):
            dA = np.zeros((n, n))
            rx = net.reactions[j]
            (a,), (b,) = rx.reactants, rx.products
            dA[a, a] -= 1.0
            dA[b, a] += 1.0
            big = np.block([[A, dA], [np.zeros((n, n)), A]])
            dx = (expm(big * t)[:n, n:] @ x0) * k[j] * math.log(10.0)  # log10 scale
            np.testing.assert_allclose(r["sx"][0, it, j], dx, rtol=1e-7, atol=1e-11)


def _fd_sllh(net, theta, kfix, ts, meas, sigma, eps=1e-4, **kw):
    g = np.zeros(net.np)
    for j in range(net.np):
        tp = theta.copy()
        tp[j] += eps / 2
        tm = theta.copy()
        tm[j] -= eps / 2
        lp = O.simulate(net, tp[None], kfix[None], ts, meas[None], sigma[None], **kw)["llh"][0]
        lm = O.simulate(net, tm[None], kfix[None], ts, meas[None], sigma[None], **kw)["llh"][0]
        g[j] = (lp - lm) / eps
    return g


@pytest.mark.parametrize("name", ["linear", "robertson"])
def test_fsa_asa_fd_agree(name):
    if name == "linear":
        net = networks.linear_chain(3)
        ts = np.linspace(0.0, 5.0, 11)
    else:
        net = networks.robertson()
        ts = networks.robertson_times()[::4]
    theta = net.theta_nominal.copy()
    kfix = np.zeros(0)
    y = O.simulate(net, theta[None], kfix[None], ts, rtol=1e-12, atol=1e-16)["y"][0]
    sigma = np.full_like(y, 0.01)
    meas = networks.add_noise(y, net.noise, 0.01, seed=7)
    kw = dict(rtol=1e-12, atol=1e-16)
    f = O.simulate(net, theta[None], kfix[None], ts, meas[None], sigma[None], sensi=1, **kw)
    a = O.simulate(net, theta[None], kfix[None], ts, meas[None], sigma[None], sensi=2,
                   quad_rtol=1e-12, quad_atol=1e-16, **kw)
    assert f["status"][0] == 0 and a["status"][0] == 0
    assert f["llh"][0] == pytest.approx(a["llh"][0], rel=1e-9)
    scale = np.abs(f["sllh"][0]).max()
    np.testing.assert_allclose(a["sllh"][0], f["sllh"][0], rtol=1e-6, atol=1e-7 * scale)
    fd = _fd_sllh(net, theta, kfix, ts, meas, sigma, **kw)
    np.testing.assert_allclose(f["sllh"][0], fd, rtol=1e-4, atol=1e-5 * scale)


def test_replicate_identity():
    """llh(2 copies of the data) == 2 llh, for FSA and ASA
    (tests/cpp/jakstat_adjoint/tests1.cpp:73-130)."""
    net = networks.robertson()
    ts = networks.robertson_times()[::5]
    theta = net.theta_nominal + 0.1
    y = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-10, atol=1e-14)["y"][0]
    meas = networks.add_noise(y, net.noise, 0.01, seed=3)
    sig = np.full_like(meas, 0.01)
    ts2 = np.repeat(ts, 2)
    meas2 = np.repeat(meas, 2, axis=0)
    sig2 = np.repeat(sig, 2, axis=0)
    for sensi in (1, 2):
        one = O.simulate(net, theta[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=sensi)
        two = O.simulate(net, theta[None], np.zeros((1, 0)), ts2, meas2[None], sig2[None], sensi=sensi)
        assert two["llh"][0] == pytest.approx(2 * one["llh"][0], rel=1e-9)
        np.testing.assert_allclose(two["sllh"][0], 2 * one["sllh"][0], rtol=1e-6,
                                   atol=1e-8 * np.abs(one["sllh"][0]).max())


def test_robertson_invariants_and_published_values():
    net = networks.robertson()
    ts = np.array([0.4, 4.0, 40.0, 400.0, 4e4])
    r = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16)
    x = r["x"][0]
    np.testing.assert_allclose(x.sum(1), 1.0, atol=1e-12)
    # reference solution of the Robertson problem at t = 40
    # (Hairer & Wanner, Solving ODEs II, sect. IV.10; SUNDIALS cvRoberts example)
    np.testing.assert_allclose(x[2], [0.7158270687, 9.185534765e-6, 0.2841637457], rtol=1e-8)
    # scipy Radau at tight tolerance on the benchmark time grid
    tg = networks.robertson_times()
    rr = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), tg, rtol=1e-12, atol=1e-16)
    om = O.OracleModel(net)
    sol = solve_ivp(lambda t, y: om.xdot(net.theta_nominal, y), (0, tg[-1]), net.x0_const,
                    method="Radau", t_eval=tg, rtol=1e-12, atol=1e-18,
                    jac=lambda t, y: om.jac(net.theta_nominal, y))
    np.testing.assert_allclose(rr["x"][0], sol.y.T, rtol=1e-6, atol=1e-14)


def test_network_against_scipy_radau():
    net = networks.network_n40()
    cs = networks.condition_set("N40")
    om = O.OracleModel(net)
    for b in (0, 7):
        x0 = np.where(net.x0_fixed >= 0, cs["kfix"][b][np.maximum(net.x0_fixed, 0)], net.x0_const)
        r = O.simulate(net, cs["theta"][b][None], cs["kfix"][b][None], cs["ts"], rtol=1e-11,
                       atol=1e-15)
        th = cs["theta"][b]
        sol = solve_ivp(lambda t, y: om.xdot(th, y), (0, cs["ts"][-1]), x0, method="Radau",
                        t_eval=cs["ts"], rtol=1e-11, atol=1e-15, jac=lambda t, y: om.jac(th, y))
        np.testing.assert_allclose(r["x"][0], sol.y.T, rtol=1e-6, atol=1e-10)


def test_chain_rule_log10():
    """sllh on the log10 scale equals p ln(10) times the gradient in p
    (src/rdata.cpp:782-830), by a finite difference in p itself."""
    net = networks.linear_chain(3)
    ts = np.linspace(0.0, 4.0, 9)
    theta = net.theta_nominal
    y = O.simulate(net, theta[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16)["y"][0]
    meas = y + 0.05
    sig = np.full_like(meas, 0.1)
    f = O.simulate(net, theta[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=1,
                   rtol=1e-12, atol=1e-16)
    p = 10.0 ** theta
    for j in range(net.np):
        h = 1e-6 * p[j]
        tp, tm = theta.copy(), theta.copy()
        tp[j], tm[j] = np.log10(p[j] + h), np.log10(p[j] - h)
        lp = O.simulate(net, tp[None], np.zeros((1, 0)), ts, meas[None], sig[None], rtol=1e-12, atol=1e-16)["llh"][0]
        lm = O.simulate(net, tm[None], np.zeros((1, 0)), ts, meas[None], sig[None], rtol=1e-12, atol=1e-16)["llh"][0]
        assert f["sllh"][0][j] == pytest.approx((lp - lm) / (2 * h) * p[j] * math.log(10.0), rel=1e-5)


def test_batch_threads_deterministic():
    """run_simulations(num_threads>1) gives the same results as one thread
    (python/tests/test_pregenerated_models.py:148-174)."""
    net = networks.robertson()
    th = networks.robertson_parameter_sets(16, 2)
    a = O.simulate(net, th, np.zeros((16, 0)), networks.robertson_times(), nthreads=1, rtol=1e-8, atol=1e-12)
    b = O.simulate(net, th, np.zeros((16, 0)), networks.robertson_times(), nthreads=4, rtol=1e-8, atol=1e-12)
    np.testing.assert_array_equal(a["x"], b["x"])


def test_preequilibration_steady_state_and_sensitivities():
    """Pre-equilibration of the open network: the oracle's steady state (CVODES
    to t = 1e6) against scipy Radau, its integrated sensitivities against the
    implicit-function solution -J^{-1} df/dp (steadystateproblem.cpp), and the
    inputs reset from kfix afterwards."""
    net = networks.network_n40_open()
    cs = networks.condition_set("N40P")
    om = O.OracleModel(net)
    th = cs["theta"][0]
    inp = net.x0_fixed >= 0
    x0 = np.where(inp, cs["kfix_preeq"][0][np.maximum(net.x0_fixed, 0)], net.x0_const)
    sol = solve_ivp(lambda t, y: om.xdot(th, y), (0, 1e6), x0, method="Radau", rtol=1e-12,
                    atol=1e-16, jac=lambda t, y: om.jac(th, y))
    xs = sol.y[:, -1]
    r = O.simulate(net, th[None], cs["kfix"][:1], np.array([0.0]), rtol=1e-10, atol=1e-14,
                   kpre=cs["kfix_preeq"][:1], sensi=1, want_sx=True)
    assert r["status"][0] == 0
    xo = r["x"][0, 0]
    np.testing.assert_allclose(xo[~inp], xs[~inp], rtol=1e-8, atol=1e-14)
    np.testing.assert_array_equal(xo[inp], cs["kfix"][0][net.x0_fixed[inp]])
    sx_ift = -np.linalg.solve(om.jac(th, xs), om.dxdotdp(th, xs)) * (10.0 ** th) * math.log(10.0)
    sx_ift[inp] = 0.0
    np.testing.assert_allclose(r["sx"][0, 0].T, sx_ift, rtol=1e-6, atol=1e-9 * np.abs(sx_ift).max())


def test_preequilibration_gradients():
    """Forward and adjoint gradients through pre-equilibration agree with each
    other and with central finite differences of llh (log10 scale)."""
    net = networks.network_n40_open()
    cs = networks.condition_set("N40P")
    b = slice(3, 4)
    kw = dict(rtol=1e-11, atol=1e-15, kpre=cs["kfix_preeq"][b])
    th = cs["theta"][b].copy()
    y = O.simulate(net, th, cs["kfix"][b], cs["ts"], **kw)["y"]
    meas = networks.add_noise(y, net.noise, 0.1, seed=8)
    sig = np.full_like(meas, 0.1)
    f = O.simulate(net, th, cs["kfix"][b], cs["ts"], meas, sig, sensi=1, **kw)
    a = O.simulate(net, th, cs["kfix"][b], cs["ts"], meas, sig, sensi=2, quad_rtol=1e-11,
                   quad_atol=1e-15, **kw)
    assert f["status"][0] == 0 and a["status"][0] == 0
    scale = np.abs(f["sllh"][0]).max()
    np.testing.assert_allclose(a["sllh"][0], f["sllh"][0], rtol=1e-6, atol=1e-7 * scale)
    eps = 1e-4
    for j in np.argsort(-np.abs(f["sllh"][0]))[:3]:
        tp, tm = th.copy(), th.copy()
        tp[0, j] += eps / 2
        tm[0, j] -= eps / 2
        lp = O.simulate(net, tp, cs["kfix"][b], cs["ts"], meas, sig, **kw)["llh"][0]
        lm = O.simulate(net, tm, cs["kfix"][b], cs["ts"], meas, sig, **kw)["llh"][0]
        assert f["sllh"][0][j] == pytest.approx((lp - lm) / eps, rel=1e-4, abs=1e-6 * scale)


def test_step_budget_counts_over_the_whole_run():
    """Solver::apply_max_num_steps (src/solver.cpp:107-119) hands CVode the
    steps left of the whole run's budget: a budget used up inside an output
    interval ends with CV_TOO_MUCH_WORK (-1), one used up exactly at an output
    time with the AmiException of the next call, AMICI_ERROR (-99,
    src/amici.cpp:156-157).  The cumulative steps to each output time come
    from successful runs on truncated time grids, so which case a budget
    falls into is known independently of the runs with the budget."""
    g = np.load(GOLD / "robertson_step_limit.npz")
    net = networks.robertson()
    steps = g["steps"]
    assert np.all(np.diff(steps) > 0)
    for m, st, bad in zip(g["budgets"], g["status"], g["first_bad"]):
        r = O.simulate(net, g["theta"], np.zeros((1, 0)), g["ts"], maxsteps=int(m))
        assert int(r["status"][0]) == st
        assert st == (-99 if m in steps else -1)
        assert bad == int((steps <= m).sum())


def test_forward_trace_matches_counters():
    """The oracle's step trace (CV_ONE_STEP with tstop at every output time)
    takes the same steps as the CV_NORMAL run whose counters the fixtures
    hold: same number of steps, function evaluations and setups, ending at the
    last output time."""
    net = networks.robertson()
    ts = networks.robertson_times()
    th = networks.robertson_parameter_sets(2, seed=9)
    for b in range(2):
        tr = O.trace_forward(net, th[b], np.zeros(0), ts, rtol=1e-8, atol=1e-12)
        c = O.simulate(net, th[b:b + 1], np.zeros((1, 0)), ts, rtol=1e-8, atol=1e-12)["counters"][0]
        assert len(tr) == c[0] and tr[-1, 3] == c[0]
        assert tr[-1, 4] == c[1] and tr[-1, 5] == c[2]
        assert tr[-1, 0] == ts[-1] and np.all(np.diff(tr[:, 0]) > 0)


def test_parameter_scales_by_finite_differences():
    """Each parameter on its own scale (ParameterScaling none / ln / log10,
    Model::set_parameter_scale): theta unscaled per parameter, and sllh the
    gradient with respect to theta on that scale, by central differences of
    llh in theta itself (linear chain, both sensitivity methods)."""
    net = networks.linear_chain(3)
    ts = np.linspace(0.0, 4.0, 9)
    p = 10.0 ** net.theta_nominal
    pscale = np.array([0, 1, 2, 0])[: net.np]
    theta = np.where(pscale == 0, p, np.where(pscale == 1, np.log(p), np.log10(p)))
    kw = dict(rtol=1e-12, atol=1e-16, pscale=pscale)
    y = O.simulate(net, theta[None], np.zeros((1, 0)), ts, **kw)["y"][0]
    ref = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16)
    np.testing.assert_allclose(y, ref["y"][0], rtol=1e-12)
    meas = y + 0.05
    sig = np.full_like(meas, 0.1)
    for sensi in (1, 2):
        f = O.simulate(net, theta[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=sensi,
                       quad_rtol=1e-12, quad_atol=1e-16, **kw)
        for j in range(net.np):
            h = 1e-6 * max(abs(theta[j]), 1.0)
            tp, tm = theta.copy(), theta.copy()
            tp[j] += h
            tm[j] -= h
            lp = O.simulate(net, tp[None], np.zeros((1, 0)), ts, meas[None], sig[None], **kw)["llh"][0]
            lm = O.simulate(net, tm[None], np.zeros((1, 0)), ts, meas[None], sig[None], **kw)["llh"][0]
            assert f["sllh"][0][j] == pytest.approx((lp - lm) / (2 * h), rel=1e-5), (sensi, j)


def test_parameter_list_selects_sensitivities():
    """Sensitivities with respect to a parameter list (Model::set_parameter_list,
    ExpData::plist; any order, a subset): sllh and sx are the columns of the
    full run in the list's order, for both methods, up to the error-control
    differences of integrating fewer sensitivities / quadratures."""
    net = networks.network_n40()
    g = np.load(GOLD / "n40_conditions16.npz")
    th, kf, ts, me, sg = g["theta"][:2], g["kfix"][:2], g["ts"], g["meas"][:2], g["sigma"][:2]
    kw = dict(rtol=1e-10, atol=1e-14, quad_rtol=1e-10, quad_atol=1e-14)
    pl = [117, 3, 64, 5]
    for sensi, key in ((1, "sllh_fsa"), (2, "sllh_asa")):
        r = O.simulate(net, th, kf, ts, me, sg, sensi=sensi, plist=pl, want_sx=True, **kw)
        assert r["sllh"].shape == (2, len(pl))
        full = g[key][:2, pl]
        assert np.abs(r["sllh"] - full).max() <= 1e-6 * np.abs(full).max()
        if sensi == 1:
            f = O.simulate(net, th, kf, ts, me, sg, sensi=1, want_sx=True, **kw)
            np.testing.assert_allclose(r["sx"], f["sx"][:, :, pl], rtol=1e-6,
                                       atol=1e-9 * np.abs(f["sx"]).max())


def test_initial_state_override():
    """ExpData::x0 replaces the model's initial state (Model::set_initial_state):
    the linear chain from a given x0 follows expm(A t) x0, and x0 equal to the
    model's own initial state reproduces the default run bit for bit."""
    net = networks.linear_chain(3)
    ts = np.array([0.0, 0.5, 2.0, 6.0])
    x0 = np.array([0.3, 1.7, 0.2, 0.9, 0.4])[: net.nx]
    r = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16,
                   x0=x0[None])
    A, _ = _linear_A(net, net.theta_nominal)
    for it, t in enumerate(ts):
        np.testing.assert_allclose(r["x"][0, it], expm(A * t) @ x0, rtol=1e-9, atol=1e-13)
    a = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-10, atol=1e-14)
    b = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-10, atol=1e-14,
                   x0=net.x0_const[None])
    assert np.array_equal(a["x"], b["x"])


def test_initial_state_sensitivities():
    """ExpData::sx0 (Model::set_initial_state_sensitivities, src/model.cpp:986-1023),
    given with respect to the scaled parameters: the linear chain's forward
    sensitivities are expm(A t) sx0 plus the parameters' own contribution
    (Van Loan, as in test_linear_states_and_fsa_closed_form), and the adjoint
    gradient with sx0 equals the forward one."""
    net = networks.linear_chain(3)
    n = net.nx
    ts = np.array([0.0, 0.5, 2.0, 6.0])
    rng = np.random.default_rng(3)
    sx0 = rng.normal(size=(net.np, n))
    r = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16,
                   sensi=1, want_sx=True, sx0=sx0[None])
    A, k = _linear_A(net, net.theta_nominal)
    x0 = net.x0_const
    for it, t in enumerate(ts):
        for j in range(net.np):
            dA = np.zeros((n, n))
            rx = net.reactions[j]
            (a,), (b,) = rx.reactants, rx.products
            dA[a, a] -= 1.0
            dA[b, a] += 1.0
            big = np.block([[A, dA], [np.zeros((n, n)), A]])
            dx = (expm(big * t)[:n, n:] @ x0) * k[j] * math.log(10.0) + expm(A * t) @ sx0[j]
            np.testing.assert_allclose(r["sx"][0, it, j], dx, rtol=1e-7, atol=1e-11)
    y = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, rtol=1e-12, atol=1e-16)["y"][0]
    meas = y + 0.05
    sig = np.full_like(meas, 0.1)
    kw = dict(rtol=1e-12, atol=1e-16, quad_rtol=1e-12, quad_atol=1e-16, sx0=sx0[None])
    f = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=1, **kw)
    a = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=2, **kw)
    z = O.simulate(net, net.theta_nominal[None], np.zeros((1, 0)), ts, meas[None], sig[None], sensi=2,
                   rtol=1e-12, atol=1e-16, quad_rtol=1e-12, quad_atol=1e-16)
    np.testing.assert_allclose(a["sllh"], f["sllh"], rtol=1e-6, atol=1e-9)
    assert np.abs(a["sllh"] - z["sllh"]).max() > 1e-3   # sx0 changes the gradient
