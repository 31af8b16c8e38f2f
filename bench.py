"""Benchmark: adjoint NLL+gradient evaluations per second of the 40-species /
120-reaction mass-action network over 1024 input conditions (rtol 1e-8,
atol 1e-12, FP64) on one MI355X, beside the plain GPU BDF (one instance per
launch, forward sensitivities) and the CVODES CPU path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config N40|N200]

One step = one batched adjoint NLL+gradient evaluation of all conditions of
the configuration (llh and the full sllh of each), inputs resident in HBM.
--config N200 measures the 200-species / 600-reaction network over 256
conditions (one instance per 256-lane workgroup) the same way; it has no
plain-GPU-BDF line (that path needs forward sensitivities, which the
workgroup engine does not provide).  --config RB measures 4096 Robertson
parameter sets with forward sensitivities.  --config MIX runs the mixed workload
(2048 conditions of three model types, pre-equilibrated where the network
has a steady state) as three concurrent batches.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "NLL+gradient evals/s at rtol 1e-8: batched adjoint vs plain GPU BDF, 1 GPU"
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
C_JY = 16  # flops of Jy + dJy/dy per observable and time point (log-normal)
TOL = dict(rtol=1e-8, atol=1e-12, quad_rtol=1e-8, quad_atol=1e-12, maxsteps=10000)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores() -> tuple:
    """Threads for the CPU baseline: every core this process may run on
    (its affinity set), limited by the CPU share the machine grants one GPU
    job when it states one in OMP_NUM_THREADS.  Returns (threads, note)."""
    aff = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    n = aff
    if share.isdigit() and 0 < int(share) < aff:
        n = int(share)
    note = (f"{n} threads: affinity set {aff} cores, os.cpu_count() {os.cpu_count()}, "
            f"OMP_NUM_THREADS {share or 'unset'} (the CPU share of this GPU job)")
    return n, note


def flop_model(counters: np.ndarray, n: int, ops: dict, nt: int, ny: int) -> tuple:
    """Algorithmic FLOPs per instance of the forward and of the backward kernel,
    from CVODES counters at rtol 1e-8 (dense-equivalent LU and solve costs):

      forward  = nfe c_f + nje c_J + nsetups (2/3 n^3 + n^2) + nni 2 n^2
                 + nt (c_y + ny C_JY + 2 ny n)
      backward = nfeB (2 n^2 + c_fpT) + nsetupsB (2/3 n^3 + n^2) + nniB 2 n^2
                 + nfQB c_fpT,   c_fpT = 2 nnz(df/dp)
    """
    c = counters.astype(np.float64)
    lu = 2.0 / 3.0 * n ** 3 + n ** 2
    sv = 2.0 * n * n
    c_fpT = 2.0 * ops["c_fp"]
    fwd = (c[:, 1] * ops["c_f"] + c[:, 3] * ops["c_J"] + c[:, 2] * lu + c[:, 4] * sv
           + nt * (ops["c_y"] + ny * C_JY + 2 * ny * n))
    bwd = c[:, 10] * (sv + c_fpT) + c[:, 11] * lu + c[:, 12] * sv + c[:, 13] * c_fpT
    return fwd, bwd


CONFIGS = {
    "N40": dict(counters=("n40_counters1024.npz", "counters_asa"), cpu_sample=256,
                plain_sample=8, traffic="n40_adjoint_hbm_traffic.json",
                workload="N40 adjoint NLL+gradient: 40 species, 120 reactions, 1024 input "
                         "conditions, 25 times, 30 log-normal observables"),
    "N200": dict(counters=("n200_conditions4.npz", "counters_bench"), cpu_sample=16,
                 plain_sample=0, traffic="n200_adjoint_ck32_hbm_traffic.json", ckpt=32,
                 workload="N200 adjoint NLL+gradient: 200 species, 600 reactions, 256 input "
                          "conditions, 40 times, 60 log-normal observables, forward "
                          "checkpoints every 32 steps with recomputation"),
}


def traffic_from_profile(name: str):
    """HBM bytes per launch of the dominant kernel from the committed
    rocprofv3 counter summary, if present."""
    p = ROOT / "profiles" / name
    if not p.exists():
        return None
    try:
        return json.loads(p.read_text())
    except ValueError:
        return None


def cpu_baseline(cs, theta, kfix, meas, sigma, sample: int, config: str):
    """The reference's CVODES (oracle/_ref, built from the reference sources)
    on the host cores: adjoint NLL+gradient of the first `sample` conditions."""
    try:
        from oracle import cvodes_oracle as O
        O.lib()
    except Exception as e:  # library not built on this machine
        log(f"cpu baseline skipped: {e}")
        return None
    cores, note = host_cores()
    t0 = time.perf_counter()
    r = O.simulate(cs["network"], theta[:sample], kfix[:sample], cs["ts"], meas[:sample],
                   sigma[:sample], sensi=2, nthreads=cores, **TOL)
    dt = time.perf_counter() - t0
    ok = int((r["status"] == 0).sum())
    return {"value": ok / dt, "unit": "evals/s", "cores": cores, "cores_note": note,
            "kind": "reference",
            "sample": f"{sample} of {theta.shape[0]} {config} conditions, adjoint NLL+gradient, "
                      f"CVODES 7.1.1 "
                      f"(reference sources) + AMICI layer restated in C, OpenMP over conditions, "
                      f"{dt:.2f} s wall, {ok}/{sample} succeeded"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", choices=sorted(CONFIGS) + ["MIX", "RB"], default="N40")
    ap.add_argument("--plain-sample", type=int, default=None,
                    help="conditions timed on the plain GPU BDF path (0 = skip; N40: 8)")
    ap.add_argument("--cpu-sample", type=int, default=None,
                    help="conditions timed on the CPU baseline (0 = skip; N40: 256, N200: 16)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a device-to-device copy of BYTES bytes 5 times (the known-size "
                         "kernel that scales the HBM counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the batched path shards by instances across GPUs "
              "with no collective on the hot path, and only the 1-GPU figure is measured.",
              file=sys.stderr)
        return 2

    import torch

    from amici_amd import codegen, networks, sim

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        return 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_calibration:
        n = args.copy_calibration // 8
        src = torch.rand(n, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 8, "launches": 5}))
        return 0
    if args.config == "MIX":
        return run_mix(args, dev)
    if args.config == "RB":
        return run_rb(args, dev)
    cfg = CONFIGS[args.config]
    if args.plain_sample is None:
        args.plain_sample = cfg["plain_sample"]
    if args.cpu_sample is None:
        args.cpu_sample = cfg["cpu_sample"]
    cs = networks.condition_set(args.config)
    net = cs["network"]
    model = sim.Model(net)
    nb, nt = cs["theta"].shape[0], len(cs["ts"])
    theta = torch.as_tensor(cs["theta"], device=dev)
    kfix = torch.as_tensor(cs["kfix"], device=dev)
    ts = torch.as_tensor(cs["ts"], device=dev)

    # synthetic data: this path's own observables at the nominal parameters,
    # with the configuration's noise (log-normal, sigma 0.1, its seed)
    states = sim.Solver(**TOL)
    out = sim.simulate_batch(model, states, theta, kfix, ts, want_x=False)
    torch.cuda.synchronize()
    if int((out["status"] != 0).sum()):
        raise RuntimeError("data generation failed")
    y = out["y"].cpu().numpy()
    meas_np = networks.add_noise(y, net.noise, cs["sigma"], seed=cs["noise_seed"])
    sigma_np = np.full_like(meas_np, cs["sigma"])
    meas = torch.as_tensor(meas_np, device=dev)
    sigma = torch.as_tensor(sigma_np, device=dev)

    adj = sim.Solver(**TOL)
    adj.set_sensitivity_order(sim.SensitivityOrder.first)
    adj.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    adj.set_checkpoint_steps(cfg.get("ckpt", 0))

    # the timed loop replays the batch captured once as a HIP graph
    # (counter reset, forward kernel, backward kernel)
    graph = sim.BatchGraph(model, adj, theta, kfix, ts, meas, sigma, want_x=False, want_y=False,
                           want_stats=False)

    def eager():
        return sim.simulate_batch(model, adj, theta, kfix, ts, meas, sigma, want_x=False,
                                  want_y=False, want_stats=False)

    for _ in range(args.warmup):
        graph.launch()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = graph.launch()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    # kernel durations: the same launches issued eagerly, timed with HIP events
    # recorded on their stream by the library
    model.lib.set_timing(True)
    model.lib.kernel_times()  # discard earlier records
    for _ in range(args.steps):
        eager()
    torch.cuda.synchronize()
    kt = model.lib.kernel_times()
    model.lib.set_timing(False)
    status = res["status"].cpu().numpy()
    sllh = res["sllh"].cpu().numpy()
    n_ok = int(((status == 0) & np.isfinite(sllh).all(axis=1)).sum())
    if n_ok != nb:
        log(f"warning: {nb - n_ok} instances failed: {np.unique(status, return_counts=True)}")
    # only successful evaluations count
    value = n_ok * args.steps / dt

    # solver counters of this path (for the record) and the CVODES counters
    # that define the algorithmic work
    stats = sim.simulate_batch(model, adj, theta, kfix, ts, meas, sigma, want_x=False,
                               want_y=False)["stats"].cpu().numpy()
    # CVODES counters at rtol 1e-8: every N40 condition; for N200 the 4
    # conditions of the fixture stand for the batch (per-instance mean x nb)
    cf, ck = cfg["counters"]
    gold = np.load(ROOT / "tests" / "golden" / cf)[ck]
    ops = codegen.op_counts(net)
    f_fwd, f_bwd = flop_model(gold, net.nx, ops, nt, net.ny)
    if len(gold) != nb:
        f_fwd = np.full(nb, f_fwd.mean())
        f_bwd = np.full(nb, f_bwd.mean())
    fwd_ms = kt["fwd_ms"] / max(kt["fwd_launches"], 1)
    bwd_ms = kt["bwd_ms"] / max(kt["bwd_launches"], 1)
    dom = "bwd" if bwd_ms >= fwd_ms else "fwd"
    dom_flops = float((f_bwd if dom == "bwd" else f_fwd).sum())
    dom_ms = bwd_ms if dom == "bwd" else fwd_ms
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    tr = traffic_from_profile(cfg["traffic"])
    # the compute roofline: MI355X's FP64 peak (78.6 TFLOP/s for the vector
    # units and for the matrix cores alike); what limits the kernel is stated
    # beside it
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                "issued": "FP64 VALU (this kernel issues no MFMA instruction)",
                "limiter": "per-instance dependent-instruction latency (one wavefront per SIMD)",
                "traffic": (tr or {}).get("bytes_per_launch") if tr else None,
                "kernel": ("amx_bwd_adj" if dom == "bwd" else "amx_fwd_adj")
                          + ("_ck" if cfg.get("ckpt") else ""),
                "kernel_ms": dom_ms, "flops_per_launch": dom_flops,
                "fwd_kernel_ms": fwd_ms, "bwd_kernel_ms": bwd_ms,
                "fwd_flops_per_launch": float(f_fwd.sum()),
                "bwd_flops_per_launch": float(f_bwd.sum()),
                "whole_step_tflops": float((f_fwd + f_bwd).sum()) / (dt / args.steps) / 1e12}

    # plain GPU BDF: one instance per launch, forward sensitivities
    plain, grad_rel = None, None
    if args.plain_sample > 0:
        plain, grad_rel = plain_gpu_bdf(model, theta, kfix, ts, meas, sigma, sllh, nb, args.plain_sample)

    cpu = cpu_baseline(cs, cs["theta"], cs["kfix"], meas_np, sigma_np, args.cpu_sample,
                       args.config) if args.cpu_sample > 0 else None
    accuracy = accuracy_gate(model, adj, dev) if args.config == "N40" else None

    line = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": cfg["workload"],
                   "instances": nb, "np": net.np, "rtol": TOL["rtol"], "atol": TOL["atol"],
                   "sensitivity": "adjoint", "checkpoint_steps": cfg.get("ckpt", 0)},
        "instances_ok": n_ok,
        "plain_gpu_bdf": plain,
        "speedup_vs_plain_gpu_bdf": (value / plain["value"]) if plain else None,
        "sllh_adjoint_vs_forward_rel": grad_rel,
        "solver_steps_mean": {"fwd": float(stats[:, 0].mean()), "bwd": float(stats[:, 8].mean()),
                              "cvodes_fwd": float(gold[:, 0].mean()),
                              "cvodes_bwd": float(gold[:, 9].mean())},
        "accuracy": accuracy,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _rel(a, ref, floor):
    """Largest |a - ref| / (|ref| + floor * max|ref| of the instance)."""
    scale = np.abs(ref).reshape(ref.shape[0], -1).max(axis=1)
    scale = scale.reshape((-1,) + (1,) * (ref.ndim - 1))
    return float(np.max(np.abs(a - ref) / (np.abs(ref) + floor * scale)))


def _normwise(a, ref):
    """Per instance max|a - ref| / max|ref| (the gradient as a vector)."""
    return np.abs(a - ref).max(axis=1) / np.abs(ref).max(axis=1)


def accuracy_gate(model, solver, dev):
    """The timed solver settings (rtol 1e-8 / atol 1e-12) on the 16 N40
    conditions of the oracle fixtures: errors against the oracle at rtol
    1e-10 beside the oracle's own errors at the timed tolerances."""
    from amici_amd import sim

    g = np.load(ROOT / "tests" / "golden" / "n40_conditions16.npz")
    c = np.load(ROOT / "tests" / "golden" / "n40_conditions16_rtol8.npz")
    out = sim.simulate_batch(model, solver, g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"])
    import torch
    torch.cuda.synchronize()
    st = out["status"].cpu().numpy()
    x, y = out["x"].cpu().numpy(), out["y"].cpu().numpy()
    llh, sllh = out["llh"].cpu().numpy(), out["sllh"].cpu().numpy()
    gpu_nw, cv_nw = _normwise(sllh, g["sllh_asa"]), _normwise(c["sllh_asa"], g["sllh_asa"])
    return {
        "conditions": int(len(st)), "succeeded": int((st == 0).sum()),
        "reference": "oracle (CVODES 7.1.1) at rtol 1e-10 / atol 1e-14",
        "gpu": {"x_rel": _rel(x, g["x"], 1e-9), "y_rel": _rel(y, g["y"], 1e-9),
                "llh_rel": float(np.max(np.abs(llh - g["llh"]) / np.abs(g["llh"]))),
                "sllh_rel_entrywise": _rel(sllh, g["sllh_asa"], 1e-7),
                "sllh_rel_normwise": float(gpu_nw.max())},
        "cvodes_rtol8": {"x_rel": _rel(c["x"], g["x"], 1e-9), "y_rel": _rel(c["y"], g["y"], 1e-9),
                         "llh_rel": float(np.max(np.abs(c["llh"] - g["llh"]) / np.abs(g["llh"]))),
                         "sllh_rel_entrywise": _rel(c["sllh_asa"], g["sllh_asa"], 1e-7),
                         "sllh_rel_normwise": float(cv_nw.max())},
    }


def run_mix(args, dev) -> int:
    """The mixed workload (networks.mix_groups: 2048 conditions, 50 / 35 / 15 %
    Robertson / open 40-species / open 200-species, the latter two
    pre-equilibrated): one step runs the three model batches at the same
    time, each on its own stream (sim.simulate_concurrent)."""
    import torch

    from amici_amd import networks, sim

    groups, total = [], 0
    for cs in networks.mix_groups():
        net = cs["network"]
        model = sim.Model(net)
        theta = torch.as_tensor(cs["theta"], device=dev)
        kfix = torch.as_tensor(cs["kfix"], device=dev)
        ts = torch.as_tensor(cs["ts"], device=dev)
        kpre = cs.get("kfix_preeq")
        kpre = torch.as_tensor(kpre, device=dev) if kpre is not None else None
        out = sim.simulate_batch(model, sim.Solver(**TOL), theta, kfix, ts, want_x=False,
                                 kfix_preeq=kpre)
        torch.cuda.synchronize()
        if int((out["status"] != 0).sum()):
            raise RuntimeError(f"data generation failed for {net.name}")
        meas_np = networks.add_noise(out["y"].cpu().numpy(), net.noise, cs["sigma"],
                                     seed=cs["noise_seed"])
        adj = sim.Solver(**TOL)
        adj.set_sensitivity_order(sim.SensitivityOrder.first)
        adj.set_sensitivity_method(sim.SensitivityMethod.adjoint)
        groups.append(dict(model=model, solver=adj, theta=theta, kfix=kfix, ts=ts,
                           meas=torch.as_tensor(meas_np, device=dev),
                           sigma=torch.full(meas_np.shape, cs["sigma"], dtype=torch.float64,
                                            device=dev),
                           kfix_preeq=kpre, want_x=False, want_y=False, want_stats=False,
                           _cs=cs, _meas=meas_np))
        total += theta.shape[0]
    run = [{k: v for k, v in g.items() if not k.startswith("_")} for g in groups]

    def step():
        return sim.simulate_concurrent(run)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ok = sum(int((r["status"] == 0).sum()) for r in res)
    # each group alone, one after another (same work, no overlap)
    alone = []
    for g in run:
        kw = dict(g)
        t1 = time.perf_counter()
        sim.simulate_batch(kw.pop("model"), kw.pop("solver"), kw.pop("theta"), kw.pop("kfix"),
                           kw.pop("ts"), **kw)
        torch.cuda.synchronize()
        alone.append(time.perf_counter() - t1)
    cpu = None
    if args.cpu_sample is None or args.cpu_sample > 0:
        cpu = mix_cpu_baseline(groups)
    line = {
        "metric": METRIC, "value": total * args.steps / dt, "unit": "evals/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "MIX adjoint NLL+gradient: 2048 conditions, 1024 Robertson, "
                               "717 open 40-species and 307 open 200-species (the last two "
                               "pre-equilibrated), three model batches on concurrent streams",
                   "instances": total, "rtol": TOL["rtol"], "atol": TOL["atol"],
                   "sensitivity": "adjoint"},
        "instances_ok": ok,
        "groups": [{"model": g["model"].network.name, "instances": int(g["theta"].shape[0]),
                    "alone_ms": a * 1e3} for g, a in zip(groups, alone)],
        "concurrency_gain": sum(alone) / (dt / args.steps),
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def run_rb(args, dev) -> int:
    """RB: 4096 Robertson parameter sets (log10 k ~ N(nominal, 0.5)), forward
    sensitivities of llh with respect to the 3 rate constants, one batch per
    step; beside it the plain GPU BDF (one launch per set) and CVODES on the
    host cores, both on a sample."""
    import torch

    from amici_amd import networks, sim

    cs = networks.condition_set("RB")
    net = cs["network"]
    model = sim.Model(net)
    nb = cs["theta"].shape[0]
    theta = torch.as_tensor(cs["theta"], device=dev)
    kfix = torch.as_tensor(cs["kfix"], device=dev)
    ts = torch.as_tensor(cs["ts"], device=dev)
    # data: the nominal parameters' observables with N(0, 0.01^2) noise (seed 2)
    nom = torch.as_tensor(np.repeat(net.theta_nominal[None, :], nb, 0), device=dev)
    out = sim.simulate_batch(model, sim.Solver(**TOL), nom, kfix, ts, want_x=False)
    torch.cuda.synchronize()
    meas_np = networks.add_noise(out["y"].cpu().numpy(), net.noise, cs["sigma"], seed=cs["noise_seed"])
    sigma_np = np.full_like(meas_np, cs["sigma"])
    meas, sigma = torch.as_tensor(meas_np, device=dev), torch.as_tensor(sigma_np, device=dev)
    fsa = sim.Solver(**TOL)
    fsa.set_sensitivity_order(sim.SensitivityOrder.first)
    fsa.set_sensitivity_method(sim.SensitivityMethod.forward)

    def step():
        return sim.simulate_batch(model, fsa, theta, kfix, ts, meas, sigma, want_x=False,
                                  want_y=False, want_stats=False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    model.lib.set_timing(True)
    model.lib.kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    kt = model.lib.kernel_times()
    model.lib.set_timing(False)
    status = res["status"].cpu().numpy()
    ok = int((status == 0).sum())
    S = args.plain_sample if args.plain_sample else 64
    sl = slice(0, S)
    sim.simulate_batch(model, fsa, theta[:1], kfix[:1], ts, meas[:1], sigma[:1], sequential=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pl = sim.simulate_batch(model, fsa, theta[sl], kfix[sl], ts, meas[sl], sigma[sl], sequential=True,
                            want_x=False, want_y=False)
    torch.cuda.synchronize()
    plain_dt = time.perf_counter() - t1
    plain = {"value": S / plain_dt, "unit": "evals/s",
             "sample": f"first {S} of {nb} parameter sets, "
                       f"{int((pl['status'] == 0).sum())}/{S} succeeded, one launch each"}
    cpu = None
    ns = args.cpu_sample if args.cpu_sample else 1024
    try:
        from oracle import cvodes_oracle as O
        O.lib()
        cores, note = host_cores()
        t2 = time.perf_counter()
        r = O.simulate(net, cs["theta"][:ns], cs["kfix"][:ns], cs["ts"], meas_np[:ns], sigma_np[:ns],
                       sensi=1, nthreads=cores, **TOL)
        cdt = time.perf_counter() - t2
        cpu = {"value": int((r["status"] == 0).sum()) / cdt, "unit": "evals/s", "cores": cores,
               "cores_note": note, "kind": "reference",
               "sample": f"{ns} of {nb} RB parameter sets, forward sensitivities, CVODES 7.1.1 "
                         f"(reference sources) + AMICI layer restated in C, {cdt:.2f} s wall, "
                         f"{int((r['status'] == 0).sum())}/{ns} succeeded"}
    except Exception as e:  # library not built on this machine
        log(f"cpu baseline skipped: {e}")
    value = nb * args.steps / dt
    line = {
        "metric": METRIC, "value": value, "unit": "evals/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "RB forward-sensitivity NLL+gradient: Robertson, 4096 parameter sets "
                               "(log10 k ~ N(nominal, 0.5)), 40 times, 2 observables",
                   "instances": nb, "np": net.np, "rtol": TOL["rtol"], "atol": TOL["atol"],
                   "sensitivity": "forward"},
        "instances_ok": ok,
        "status_counts": {str(int(k)): int(v) for k, v in zip(*np.unique(status, return_counts=True))},
        "fwd_kernel_ms": kt["fwd_ms"] / max(kt["fwd_launches"], 1),
        "plain_gpu_bdf": plain, "speedup_vs_plain_gpu_bdf": value / plain["value"],
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def mix_cpu_baseline(groups):
    """The reference's CVODES on the host cores over a sample of each group,
    combined into the rate for the whole mix."""
    try:
        from oracle import cvodes_oracle as O
        O.lib()
    except Exception as e:  # library not built on this machine
        log(f"cpu baseline skipped: {e}")
        return None
    cores, note = host_cores()
    samples = {"robertson": 128, "n40open": 32, "n200open": 16}
    t_total, parts = 0.0, []
    for g in groups:
        cs = g["_cs"]
        name = cs["network"].name
        k = samples[name]
        # the 200-species group is timed without its pre-equilibration (with it, CVODES
        # integrates 805 forward sensitivities to t = 1e6: minutes per condition), so
        # the CPU rate of the mix is an upper bound
        kw = ({"kpre": cs["kfix_preeq"][:k]} if "kfix_preeq" in cs and name != "n200open"
              else {})
        t0 = time.perf_counter()
        O.simulate(cs["network"], cs["theta"][:k], cs["kfix"][:k], cs["ts"], g["_meas"][:k],
                   np.full(g["_meas"][:k].shape, cs["sigma"]), sensi=2, nthreads=cores, **TOL, **kw)
        rate = k / (time.perf_counter() - t0)
        t_total += cs["theta"].shape[0] / rate
        parts.append(f"{name} {k} conditions {rate:.1f} evals/s")
    total = sum(g["theta"].shape[0] for g in groups)
    return {"value": total / t_total, "unit": "evals/s", "cores": cores, "cores_note": note,
            "kind": "reference",
            "sample": "CVODES 7.1.1 (reference sources) + AMICI layer restated in C, OpenMP over "
                      "conditions, adjoint with pre-equilibration for the open 40-species group "
                      "(the open 200-species group without it: an upper bound on the CPU rate); "
                      "per group: " + "; ".join(parts) + "; rate of the whole mix = "
                      "2048 / sum(count / rate)"}


def plain_gpu_bdf(model, theta, kfix, ts, meas, sigma, sllh, nb, S):
    """The same integrator, one instance per kernel launch, forward
    sensitivities, on the first S conditions; also returns the largest
    relative difference between its sllh and the batched adjoint's."""
    import torch

    from amici_amd import sim

    fsa = sim.Solver(**TOL)
    fsa.set_sensitivity_order(sim.SensitivityOrder.first)
    fsa.set_sensitivity_method(sim.SensitivityMethod.forward)
    sl = slice(0, S)
    sim.simulate_batch(model, fsa, theta[:1], kfix[:1], ts, meas[:1], sigma[:1], sequential=True,
                       want_x=False, want_y=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pl = sim.simulate_batch(model, fsa, theta[sl], kfix[sl], ts, meas[sl], sigma[sl],
                            sequential=True, want_x=False, want_y=False)
    torch.cuda.synchronize()
    plain_dt = time.perf_counter() - t0
    plain_ok = int((pl["status"] == 0).sum())
    plain = {"value": S / plain_dt, "unit": "evals/s",
             "sample": f"first {S} of {nb} conditions, {plain_ok}/{S} succeeded, one launch each"}
    # agreement of the two gradients on the sample (adjoint vs forward)
    a_s = sllh[:S]
    f_s = pl["sllh"].cpu().numpy()
    grad_rel = float(np.max(np.abs(a_s - f_s) / (np.abs(f_s) + 1e-7 * np.abs(f_s).max())))
    return plain, grad_rel


if __name__ == "__main__":
    sys.exit(main())
