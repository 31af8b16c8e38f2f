"""Differentiable negative log-likelihood of a batch of conditions.

``nll(theta, model, kfix, ts, meas, sigma)`` returns the per-condition NLL
(-llh, llh as in src/model.cpp:1194-1210) as a tensor on the GPU; its backward
is the batched adjoint sweep (the gradient is -sllh, sllh from
src/backwardproblem.cpp:27-131 and src/rdata.cpp:514-583 with the log10 chain
rule of src/rdata.cpp:782-830).  The forward pass already runs the adjoint, so
backward costs nothing extra; conditions whose simulation fails give NaN NLL
and gradient.  With ``graph=True`` the batch is captured once per shape and
solver setting as a HIP graph (sim.BatchGraph) and replayed on later calls,
the new inputs copied into the graph's buffers.
"""
from __future__ import annotations

import math

import torch

from . import sim


def _graph_for(model, solver, theta, kfix, ts, meas, sigma) -> sim.BatchGraph:
    """The model's cached graph for this solver setting and input shapes
    (captured on first use), with the inputs copied in."""
    x0 = model._x0
    key = (repr(sorted(vars(solver).items())), tuple(theta.shape), tuple(torch.as_tensor(ts).shape),
           tuple(torch.as_tensor(meas).shape), tuple(model.get_parameter_scale()),
           None if x0 is None else tuple(x0),
           None if model._sx0 is None else model._sx0.tobytes(), model._t0)
    cache = model.__dict__.setdefault("_nll_graphs", {})
    g = cache.get(key)
    if g is None:
        g = sim.BatchGraph(model, solver, theta, kfix, ts, meas, sigma, want_x=False,
                           want_y=False, want_stats=False)
        cache[key] = g
    else:
        g.set_inputs(theta=theta, kfix=kfix, ts=ts, meas=meas, sigma=sigma)
    return g


class NLL(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, model, solver, kfix, ts, meas, sigma, graph=False):
        th = theta.detach().to(torch.float64)
        if graph:
            out = {k: (v.clone() if v is not None else None)
                   for k, v in _graph_for(model, solver, th, kfix, ts, meas, sigma).launch().items()}
        else:
            out = sim.simulate_batch(model, solver, th, kfix, ts, meas, sigma, want_x=False,
                                     want_y=False, want_stats=False)
        ok = (out["status"] == 0).unsqueeze(1)
        sllh = torch.where(ok, out["sllh"], torch.full_like(out["sllh"], math.nan))
        ctx.save_for_backward(sllh)
        ctx.status = out["status"]
        return torch.where(ok.squeeze(1), -out["llh"], torch.full_like(out["llh"], math.nan))

    @staticmethod
    def backward(ctx, grad):
        (sllh,) = ctx.saved_tensors
        return -grad.unsqueeze(1) * sllh, None, None, None, None, None, None, None


def adjoint_solver(**kw) -> sim.Solver:
    s = sim.Solver(**kw)
    s.set_sensitivity_order(sim.SensitivityOrder.first)
    s.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    return s


def nll(theta: torch.Tensor, model: sim.Model, kfix, ts, meas, sigma,
        solver: sim.Solver | None = None, graph: bool = False) -> torch.Tensor:
    """Per-condition NLL [nb] of the free parameters ``theta`` [nb, np] (a
    device tensor, on the model's parameter scales, log10 by default);
    differentiable w.r.t. ``theta``, which needs the sensitivities of every
    parameter (the model's parameter list in its default order).  ``graph``
    replays a captured HIP graph of the batch (one per shape and solver
    setting)."""
    if solver is None:
        solver = adjoint_solver()
    elif solver.effective_method() != sim.SensitivityMethod.adjoint:
        raise ValueError("nll needs a solver with first-order adjoint sensitivities")
    if theta.dim() != 2 or theta.shape[1] != model.np:
        raise ValueError(f"theta must be [nb, {model.np}]")
    if model.get_parameter_list() != list(range(model.np)):
        raise ValueError("nll differentiates with respect to every parameter: reset the "
                         "model's parameter list (Model.set_parameter_list(range(np)))")
    return NLL.apply(theta, model, solver, kfix, ts, meas, sigma, graph)
