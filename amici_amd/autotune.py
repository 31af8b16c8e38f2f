"""Launch-shape autotuner for the batched kernels (no counterpart in the
reference, whose CVODES runs one condition per OpenMP thread).

Two launch-shape parameters of a model module are fixed when it is compiled:

* the team: lanes per instance, one lane per state -- a 64-lane wavefront for
  models with up to 64 states, a workgroup of 128 or 256 lanes above;
* the register budget: resident waves per SIMD, 1, 2 or 4 (512, 256 or 128
  registers per lane).  A smaller budget spills more registers to scratch;
  more resident waves hide latency once a batch has more instances than the
  device has SIMDs.

``tune`` builds the candidate variants of a model (``codegen.compile_model``
with ``team`` / ``waves``), times each on the caller's own batch with the
library's kernel timers (HIP events recorded on the launch stream around the
forward and backward kernels), and returns the fastest.  The choice is cached
per device, batch shape and solver setting in ``autotune.json`` beside the
default module, so later calls with the same shape load it without timing.
"""
from __future__ import annotations

import json
from dataclasses import dataclass
from pathlib import Path

from . import codegen, sim
from .networks import ReactionNetwork


def candidates(nx: int) -> list[dict]:
    """Launch shapes worth timing for a model with ``nx`` states."""
    if nx <= 16:
        return [dict(team=64, waves=w) for w in (1, 2, 4)]
    if nx <= 64:
        return [dict(team=64, waves=w) for w in (1, 2)]
    return [dict(team=t, waves=1) for t in (128, 256) if t >= nx]


def label(c: dict) -> str:
    return f"t{c['team']}_w{c['waves']}"


def cache_path(net: ReactionNetwork) -> Path:
    return codegen.model_dir(net.name) / "autotune.json"


def cache_key(device: str, nb: int, nt: int, solver: sim.Solver) -> str:
    return (f"{device}|nb={nb}|nt={nt}|sensi={solver.effective_method()}"
            f"|interp={solver.interp}|ckpt={solver.ckpt_steps}")


@dataclass
class TuneResult:
    team: int
    waves: int
    times_ms: dict  # label -> kernel milliseconds per batch
    cached: bool

    def model(self, net: ReactionNetwork) -> sim.Model:
        return sim.Model(net, team=self.team, waves=self.waves)


def time_variant(model: sim.Model, solver: sim.Solver, batch: dict, reps: int,
                 warmup: int) -> float:
    """Kernel milliseconds per batch (forward + backward kernels, HIP events)."""
    import torch

    def run():
        return sim.simulate_batch(model, solver, **batch, want_x=False, want_y=False,
                                  want_stats=False)

    for _ in range(warmup):
        run()
    torch.cuda.synchronize()
    model.lib.set_timing(True)
    try:
        model.lib.kernel_times()  # drop earlier records
        for _ in range(reps):
            run()
        torch.cuda.synchronize()
        kt = model.lib.kernel_times()
    finally:
        model.lib.set_timing(False)
    if kt["fwd_launches"] != reps:
        raise RuntimeError(f"timed {kt['fwd_launches']} launches, expected {reps}")
    return (kt["fwd_ms"] + kt["bwd_ms"]) / reps


def tune(net: ReactionNetwork, solver: sim.Solver, theta, kfix, ts, meas=None, sigma=None, *,
         reps: int = 3, warmup: int = 1, use_cache: bool = True, device: str | None = None,
         cands: list[dict] | None = None, timer=None, loader=None) -> TuneResult:
    """Fastest launch shape of ``net`` for this batch (the arguments of
    ``sim.simulate_batch``) and solver setting.  ``timer(model, solver,
    batch, reps, warmup)`` and ``loader(candidate)`` default to the kernel
    timers and ``sim.Model``; the cache is read and written unless
    ``use_cache`` is false."""
    import numpy as np

    if device is None:
        import torch
        device = torch.cuda.get_device_name()
    nb = int(np.shape(theta)[0])
    nt = int(np.shape(ts)[-1])
    key = cache_key(device, nb, nt, solver)
    path = cache_path(net)
    cache = json.loads(path.read_text()) if path.exists() else {}
    if use_cache and key in cache:
        r = cache[key]
        return TuneResult(r["team"], r["waves"], r["times_ms"], cached=True)
    timer = timer or time_variant
    loader = loader or (lambda c: sim.Model(net, team=c["team"], waves=c["waves"]))
    cands = cands or candidates(net.nx)
    batch = dict(theta=theta, kfix=kfix, ts=ts, meas=meas, sigma=sigma)
    times = {label(c): float(timer(loader(c), solver, batch, reps, warmup)) for c in cands}
    best = min(cands, key=lambda c: times[label(c)])
    cache[key] = dict(team=best["team"], waves=best["waves"], times_ms=times)
    if use_cache:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(cache, indent=1, sort_keys=True))
    return TuneResult(best["team"], best["waves"], times, cached=False)
