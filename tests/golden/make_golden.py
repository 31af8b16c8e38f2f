"""Writes the golden fixtures under tests/golden/ from the CVODES oracle.

The reference's own golden results (tests/cpp/expected_results_py.h5) are a
Git LFS pointer in the reference checkout, so the fixtures are produced by the
oracle: SUNDIALS CVODES 7.1.1 from the reference sources, driven by the C
restatement of AMICI's forward/backward problem (oracle/cvodes_harness.c), at
rtol 1e-10 / atol 1e-14 for values and at rtol 1e-8 / atol 1e-12 for the
solver counters that define the algorithmic work of one evaluation.

Each fixture holds the inputs (theta, kfix, ts, meas, sigma) together with
the expected outputs, so the GPU parity tests need neither the reference nor
the oracle library at run time.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from amici_amd import networks  # noqa: E402
from oracle import cvodes_oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
TIGHT = dict(rtol=1e-10, atol=1e-14, quad_rtol=1e-10, quad_atol=1e-14)
BENCH = dict(rtol=1e-8, atol=1e-12, quad_rtol=1e-8, quad_atol=1e-12)


def measurements(name: str, nb: int | None = None):
    """Synthetic data of a configuration: the oracle's observables at the
    nominal parameters (per-condition inputs kept), plus the configuration's
    noise with its seed."""
    cs = networks.condition_set(name)
    net = cs["network"]
    theta = cs["theta"][:nb]
    kfix = cs["kfix"][:nb]
    nom = np.repeat(net.theta_nominal[None, :], theta.shape[0], 0)
    y = O.simulate(net, nom, kfix, cs["ts"], nthreads=8, **TIGHT)["y"]
    meas = networks.add_noise(y, net.noise, cs["sigma"], seed=cs["noise_seed"])
    sigma = np.full_like(meas, cs["sigma"])
    return cs, theta, kfix, meas, sigma


def run(net, theta, kfix, ts, meas, sigma, sensi, opts, **kw):
    r = O.simulate(net, theta, kfix, ts, meas, sigma, sensi=sensi, nthreads=8, **opts, **kw)
    assert (r["status"] == 0).all(), r["status"]
    return r


def main() -> None:
    # R: one nominal Robertson instance, states + FSA + ASA
    cs, theta, kfix, meas, sigma = measurements("R")
    net, ts = cs["network"], cs["ts"]
    f = run(net, theta, kfix, ts, meas, sigma, 1, TIGHT, want_sx=True)
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT)
    c0 = run(net, theta, kfix, ts, meas, sigma, 0, BENCH)["counters"]
    c1 = run(net, theta, kfix, ts, meas, sigma, 1, BENCH)["counters"]
    c2 = run(net, theta, kfix, ts, meas, sigma, 2, BENCH)["counters"]
    np.savez_compressed(OUT / "robertson_nominal.npz", theta=theta, kfix=kfix, ts=ts, meas=meas,
                        sigma=sigma, x=f["x"], y=f["y"], sx=f["sx"], llh=f["llh"],
                        sllh_fsa=f["sllh"], sllh_asa=a["sllh"], counters_states=c0,
                        counters_fsa=c1, counters_asa=c2)

    # RB: the first 64 of the 4096 Robertson parameter sets
    cs, theta, kfix, meas, sigma = measurements("RB", 64)
    f = run(net, theta, kfix, ts, meas, sigma, 1, TIGHT)
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT)
    np.savez_compressed(OUT / "robertson_batch64.npz", theta=theta, kfix=kfix, ts=ts, meas=meas,
                        sigma=sigma, x=f["x"], y=f["y"], llh=f["llh"], sllh_fsa=f["sllh"],
                        sllh_asa=a["sllh"])

    # N40: 16 of the 1024 input conditions
    cs, theta, kfix, meas, sigma = measurements("N40", 16)
    net, ts = cs["network"], cs["ts"]
    f = run(net, theta, kfix, ts, meas, sigma, 1, TIGHT)
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT)
    c2 = run(net, theta, kfix, ts, meas, sigma, 2, BENCH)["counters"]
    c1 = run(net, theta, kfix, ts, meas, sigma, 1, BENCH)["counters"]
    np.savez_compressed(OUT / "n40_conditions16.npz", theta=theta, kfix=kfix, ts=ts, meas=meas,
                        sigma=sigma, x=f["x"], y=f["y"], llh=f["llh"], sllh_fsa=f["sllh"],
                        sllh_asa=a["sllh"], counters_fsa=c1, counters_asa=c2)

    # edge cases on Robertson: missing data (NaN), a repeated output time,
    # output at t0 only, and a measurement-free instance
    net = networks.robertson()
    ts_e = np.array([0.0, 0.0, 1e-3, 0.5, 0.5, 10.0, 1e3])
    th = networks.robertson_parameter_sets(3, seed=9)
    y = run(net, th, np.zeros((3, 0)), ts_e, None, None, 0, TIGHT)["y"]
    meas_e = networks.add_noise(y, net.noise, 0.01, seed=9)
    meas_e[0, 2, 1] = np.nan
    meas_e[1, :, 0] = np.nan
    meas_e[2, :, :] = np.nan
    sig_e = np.full_like(meas_e, 0.02)
    f = run(net, th, np.zeros((3, 0)), ts_e, meas_e, sig_e, 1, TIGHT)
    a = run(net, th, np.zeros((3, 0)), ts_e, meas_e, sig_e, 2, TIGHT)
    np.savez_compressed(OUT / "robertson_edge.npz", theta=th, kfix=np.zeros((3, 0)), ts=ts_e,
                        meas=meas_e, sigma=sig_e, x=f["x"], y=f["y"], llh=f["llh"],
                        sllh_fsa=f["sllh"], sllh_asa=a["sllh"])

    # linear chain: closed-form-pinned small model for FSA / ASA cross-checks
    net = networks.linear_chain(3)
    ts_l = np.linspace(0.0, 5.0, 11)
    th = np.stack([net.theta_nominal, net.theta_nominal + 0.3])
    y = run(net, th, np.zeros((2, 0)), ts_l, None, None, 0, TIGHT)["y"]
    meas_l = networks.add_noise(y, net.noise, 0.05, seed=5)
    sig_l = np.full_like(meas_l, 0.05)
    f = run(net, th, np.zeros((2, 0)), ts_l, meas_l, sig_l, 1, TIGHT, want_sx=True)
    a = run(net, th, np.zeros((2, 0)), ts_l, meas_l, sig_l, 2, TIGHT)
    np.savez_compressed(OUT / "linear_chain3.npz", theta=th, kfix=np.zeros((2, 0)), ts=ts_l,
                        meas=meas_l, sigma=sig_l, x=f["x"], y=f["y"], sx=f["sx"], llh=f["llh"],
                        sllh_fsa=f["sllh"], sllh_asa=a["sllh"])
    n40_counters()
    n40_bench_accuracy()
    robertson_mix()
    n40_preeq()
    for p in sorted(OUT.glob("*.npz")):
        print(p.name, p.stat().st_size)


def n40_preeq() -> None:
    """16 conditions of the pre-equilibrated open 40-species network: steady
    state at input level 1 (CVODES integrated to t = 1e6 with forward
    sensitivities), inputs reset to the condition's levels, then 25 outputs;
    forward and adjoint gradients through the pre-equilibration."""
    cs = networks.condition_set("N40P")
    net, ts = cs["network"], cs["ts"]
    nb = 16
    theta, kfix, kpre = cs["theta"][:nb], cs["kfix"][:nb], cs["kfix_preeq"][:nb]
    y = O.simulate(net, theta, kfix, ts, nthreads=8, kpre=kpre, **TIGHT)["y"]
    meas = networks.add_noise(y, net.noise, cs["sigma"], seed=cs["noise_seed"])
    sigma = np.full_like(meas, cs["sigma"])
    f = run(net, theta, kfix, ts, meas, sigma, 1, TIGHT, kpre=kpre, want_sx=True)
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT, kpre=kpre)
    np.savez_compressed(OUT / "n40open_preeq16.npz", theta=theta, kfix=kfix, kfix_preeq=kpre,
                        ts=ts, meas=meas, sigma=sigma, x=f["x"], y=f["y"], llh=f["llh"],
                        sllh_fsa=f["sllh"], sllh_asa=a["sllh"], sx0=f["sx"][:2, 0])


def n40_bench_accuracy() -> None:
    """The 16 N40 conditions of n40_conditions16.npz solved by the oracle at
    the benchmark's tolerances (rtol 1e-8 / atol 1e-12): CVODES's own error
    at those settings, against which the timed path's error is gated."""
    g = np.load(OUT / "n40_conditions16.npz")
    net = networks.network_n40()
    args = (net, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"])
    f = run(*args, 1, BENCH)
    a = run(*args, 2, BENCH)
    np.savez_compressed(OUT / "n40_conditions16_rtol8.npz", x=a["x"], y=a["y"], llh=a["llh"],
                        sllh_asa=a["sllh"], sllh_fsa=f["sllh"])


def robertson_mix() -> None:
    """The 1024 Robertson conditions of the mixed workload (parameter sets of
    seed 5, log10 k ~ N(nominal, 0.5)): data are each set's own observables
    (oracle, rtol 1e-10) with N(0, 0.01^2) noise (seed 5).  Adjoint llh and
    sllh at rtol 1e-10 / atol 1e-14 (where the oracle's backward problem of
    one condition fails its error test: status kept) and at the benchmark's
    rtol 1e-8 / atol 1e-12 (all succeed), plus the rtol-1e-8 solver
    counters."""
    net = networks.robertson()
    g = networks.mix_groups()[0]
    theta, kfix, ts = g["theta"], g["kfix"], g["ts"]
    y = O.simulate(net, theta, kfix, ts, nthreads=8, **TIGHT)["y"]
    meas = networks.add_noise(y, net.noise, g["sigma"], seed=g["noise_seed"])
    sigma = np.full_like(meas, g["sigma"])
    t = O.simulate(net, theta, kfix, ts, meas, sigma, sensi=2, nthreads=8, **TIGHT)
    b = run(net, theta, kfix, ts, meas, sigma, 2, BENCH)
    np.savez_compressed(OUT / "robertson_mix1024.npz", theta=theta, ts=ts, meas=meas,
                        sigma=g["sigma"], llh=t["llh"], sllh_asa=t["sllh"], status=t["status"],
                        llh_rtol8=b["llh"], sllh_asa_rtol8=b["sllh"],
                        counters_rtol8=b["counters"].astype(np.int32))


def n40_counters() -> None:
    """CVODES counters of the adjoint NLL+gradient of all 1024 N40 conditions
    at rtol 1e-8 / atol 1e-12: the algorithmic work of the benchmark."""
    cs, theta, kfix, meas, sigma = measurements("N40")
    r = run(cs["network"], theta, kfix, cs["ts"], meas, sigma, 2, BENCH)
    np.savez_compressed(OUT / "n40_counters1024.npz", counters_asa=r["counters"].astype(np.int32),
                        llh=r["llh"])


def n40_stald() -> None:
    """Forward solver counters of all 1024 N40 conditions (states only, rtol
    1e-8 / atol 1e-12) with stability-limit detection on (AMICI's default)
    and off: steps and order reductions by the detector.  The inputs are
    networks.condition_set("N40"), identified by the sum of its theta."""
    cs = networks.condition_set("N40")
    net, theta, kfix, ts = cs["network"], cs["theta"], cs["kfix"], cs["ts"]
    on = O.simulate(net, theta, kfix, ts, nthreads=8, stldet=1, **BENCH)
    off = O.simulate(net, theta, kfix, ts, nthreads=8, stldet=0, **BENCH)
    assert (on["status"] == 0).all() and (off["status"] == 0).all()
    np.savez_compressed(OUT / "n40_stald1024.npz", theta_sum=theta.sum(),
                        counters_on=on["counters"].astype(np.int32),
                        counters_off=off["counters"].astype(np.int32))


def n40_plist_pscale() -> None:
    """4 N40 conditions with a parameter list (5 parameters, not in order),
    mixed parameter scales (parameter j on scale j mod 3: none, ln, log10;
    theta holds the nominal log10 values converted to those scales) and
    initial states 10 % off the model's (ExpData::x0): states, llh and the
    forward and adjoint gradients at rtol 1e-10, plus the forward sx."""
    g = np.load(OUT / "n40_conditions16.npz")
    net = networks.network_n40()
    nb = 4
    p = 10.0 ** g["theta"][:nb]
    pscale = np.arange(net.np) % 3
    theta = np.where(pscale == 0, p, np.where(pscale == 1, np.log(p), np.log10(p)))
    plist = np.array([117, 3, 64, 5, 40])
    kfix = g["kfix"][:nb]
    x0 = np.where(np.asarray(net.x0_fixed)[None] >= 0,
                  np.take_along_axis(kfix, np.maximum(np.asarray(net.x0_fixed), 0)[None].repeat(nb, 0), 1),
                  np.asarray(net.x0_const)[None])
    x0 = x0 * (1.0 + 0.1 * np.sin(np.arange(net.nx)))[None]
    args = (net, theta, kfix, g["ts"], g["meas"][:nb], g["sigma"][:nb])
    kw = dict(plist=plist, pscale=pscale, x0=x0)
    f = run(*args, 1, TIGHT, want_sx=True, **kw)
    a = run(*args, 2, TIGHT, **kw)
    np.savez_compressed(OUT / "n40_plist_pscale4.npz", theta=theta, kfix=kfix, ts=g["ts"],
                        meas=g["meas"][:nb], sigma=g["sigma"][:nb], plist=plist, pscale=pscale,
                        x0=x0, x=a["x"], y=a["y"], llh=a["llh"], sllh_fsa=f["sllh"],
                        sllh_asa=a["sllh"], sx=f["sx"])


def robertson_sx0() -> None:
    """8 Robertson conditions with caller-given initial state sensitivities
    (ExpData::sx0, with respect to the log10 parameters) for a parameter list
    [2, 0]: states, llh, the forward and adjoint gradients and sx at rtol
    1e-10."""
    g = np.load(OUT / "robertson_batch64.npz")
    net = networks.robertson()
    nb = 8
    plist = np.array([2, 0])
    sx0 = np.random.default_rng(11).normal(scale=0.1, size=(nb, len(plist), net.nx))
    args = (net, g["theta"][:nb], g["kfix"][:nb], g["ts"], g["meas"][:nb], g["sigma"][:nb])
    f = run(*args, 1, TIGHT, want_sx=True, plist=plist, sx0=sx0)
    a = run(*args, 2, TIGHT, plist=plist, sx0=sx0)
    np.savez_compressed(OUT / "robertson_sx0.npz", theta=g["theta"][:nb], kfix=g["kfix"][:nb],
                        ts=g["ts"], meas=g["meas"][:nb], sigma=g["sigma"][:nb], plist=plist,
                        sx0=sx0, x=a["x"], y=a["y"], llh=a["llh"], sllh_fsa=f["sllh"],
                        sllh_asa=a["sllh"], sx=f["sx"])


def robertson_noise() -> None:
    """16 Robertson conditions (the first parameter sets of
    robertson_batch64) with each of the noise models log10-normal, Laplace,
    log-Laplace and log10-Laplace: measurements drawn from the model around the
    oracle's observables (sigma 0.1, seed 17), then states, llh and the
    forward and adjoint gradients at rtol 1e-10.  On the log scales the points
    where the observable is below 1e-4 are missing: there (species A decays to
    1e-10) the absolute tolerance 1e-14 leaves a relative error of the
    observable that the log scale turns into gradient errors far above the
    parity bar, for CVODES's forward and adjoint gradients alike."""
    g = np.load(OUT / "robertson_batch64.npz")
    nb = 16
    base = networks.robertson()
    theta, kfix, ts = g["theta"][:nb], g["kfix"][:nb], g["ts"]
    sigma = np.full((nb, len(ts), base.ny), 0.1)
    y = run(base, theta, kfix, ts, None, sigma, 0, TIGHT)["y"]
    for noise in (networks.NOISE_LOG10NORMAL, networks.NOISE_LAPLACE, networks.NOISE_LOGLAPLACE,
                  networks.NOISE_LOG10LAPLACE):
        net = networks.with_noise(base, noise)
        meas = networks.add_noise(y, noise, 0.1, seed=17)
        if networks.noise_scale(noise):
            meas = np.where(y < 1e-4, np.nan, meas)
        args = (net, theta, kfix, ts, meas, sigma)
        f = run(*args, 1, TIGHT)
        a = run(*args, 2, TIGHT)
        np.savez_compressed(OUT / f"robertson_noise{noise}.npz", noise=noise, theta=theta,
                            kfix=kfix, ts=ts, meas=meas, sigma=sigma, x=a["x"], y=a["y"],
                            llh=a["llh"], sllh_fsa=f["sllh"], sllh_asa=a["sllh"])


def n200() -> None:
    """4 conditions of the 200-species network: states, observables, llh and
    the adjoint gradient at rtol 1e-10 / atol 1e-14, plus the rtol-1e-8
    counters of the adjoint evaluation (the algorithmic work per condition)."""
    cs, theta, kfix, meas, sigma = measurements("N200", 4)
    net, ts = cs["network"], cs["ts"]
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT)
    c = run(net, theta, kfix, ts, meas, sigma, 2, BENCH)
    np.savez_compressed(OUT / "n200_conditions4.npz", theta=theta, kfix=kfix, ts=ts, meas=meas,
                        sigma=sigma, x=a["x"], y=a["y"], llh=a["llh"], sllh_asa=a["sllh"],
                        counters_asa=a["counters"].astype(np.int32),
                        counters_bench=c["counters"].astype(np.int32))


def n200_preeq() -> None:
    """4 conditions of the pre-equilibrated open 200-species network: steady
    state at input level 1 (CVODES integrated to t = 1e6, with forward
    sensitivities for the adjoint's steady-state term), inputs reset to the
    condition's levels, then 40 outputs; states and adjoint gradients."""
    cs = networks.condition_set("N200P")
    net, ts = cs["network"], cs["ts"]
    nb = 4
    theta, kfix, kpre = cs["theta"][:nb], cs["kfix"][:nb], cs["kfix_preeq"][:nb]
    y = O.simulate(net, theta, kfix, ts, nthreads=4, kpre=kpre, **TIGHT)["y"]
    meas = networks.add_noise(y, net.noise, cs["sigma"], seed=cs["noise_seed"])
    sigma = np.full_like(meas, cs["sigma"])
    s = run(net, theta, kfix, ts, meas, sigma, 0, TIGHT, kpre=kpre)
    a = run(net, theta, kfix, ts, meas, sigma, 2, TIGHT, kpre=kpre)
    np.savez_compressed(OUT / "n200open_preeq4.npz", theta=theta, kfix=kfix, kfix_preeq=kpre,
                        ts=ts, meas=meas, sigma=sigma, x=s["x"], y=s["y"], llh=s["llh"],
                        sllh_asa=a["sllh"])


def step_limit() -> None:
    """Robertson at the nominal parameters with a forward step budget
    (Solver::set_max_steps): the budget counts over the whole run
    (Solver::apply_max_num_steps, src/solver.cpp:107-119).  Budgets that run
    out inside an output interval give AMICI_TOO_MUCH_WORK; one used up
    exactly at an output time gives AMICI_ERROR.  Stores per budget the status
    and the index of the first output time not reached, and the cumulative
    steps to each output time."""
    net = networks.robertson()
    theta = net.theta_nominal[None, :]
    ts = networks.robertson_times()
    k0 = np.zeros((1, 0))
    steps = np.array([O.simulate(net, theta, k0, ts[:k], **TIGHT)["counters"][0, 0]
                      for k in range(1, len(ts) + 1)])
    budgets = np.array([1, int(steps[10]) - 1, int(steps[10]), 500, int(steps[20]) + 1])
    status, first_bad = [], []
    for m in budgets:
        r = O.simulate(net, theta, k0, ts, maxsteps=int(m), **TIGHT)
        status.append(int(r["status"][0]))
        first_bad.append(int((steps <= m).sum()))
    np.savez_compressed(OUT / "robertson_step_limit.npz", theta=theta, ts=ts, steps=steps,
                        budgets=budgets, status=np.array(status), first_bad=np.array(first_bad))


if __name__ == "__main__":
    if sys.argv[1:] == ["n40_counters"]:
        n40_counters()
    elif sys.argv[1:] == ["n40_stald"]:
        n40_stald()
    elif sys.argv[1:] == ["n40_plist_pscale"]:
        n40_plist_pscale()
    elif sys.argv[1:] == ["robertson_sx0"]:
        robertson_sx0()
    elif sys.argv[1:] == ["n40_bench_accuracy"]:
        n40_bench_accuracy()
    elif sys.argv[1:] == ["robertson_mix"]:
        robertson_mix()
    elif sys.argv[1:] == ["step_limit"]:
        step_limit()
    elif sys.argv[1:] == ["robertson_noise"]:
        robertson_noise()
    elif sys.argv[1:] == ["n200"]:
        n200()
    elif sys.argv[1:] == ["n200_preeq"]:
        n200_preeq()
    elif sys.argv[1:] == ["n40_preeq"]:
        n40_preeq()
    else:
        main()
