"""GPU: a batch captured as a HIP graph (amx_graph_create, sim.BatchGraph)
replays the launch sequence of amx_run_batch.  Each replay must equal the
eager result bit for bit, also with eager work and synchronisations between
the replays, and after inputs are rewritten in place."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

from amici_amd import networks

GOLD = Path(__file__).resolve().parent / "golden"
BENCH = dict(rtol=1e-8, atol=1e-12, quad_rtol=1e-8, quad_atol=1e-12, maxsteps=10000)

pytestmark = pytest.mark.gpu


def _np(out):
    return {k: v.cpu().numpy() for k, v in out.items() if v is not None and k != "_keepalive"}


@pytest.mark.parametrize("name,fixture,meth,ckpt", [("robertson", "robertson_batch64", 2, 0),
                                                    ("robertson", "robertson_batch64", 1, 0),
                                                    ("n40", "n40_conditions16", 2, 0),
                                                    ("n40", "n40_conditions16", 0, 0),
                                                    ("n40", "n40_conditions16", 2, 32)])
def test_graph_replays_equal_eager(gpu, name, fixture, meth, ckpt):
    import torch
    from amici_amd import sim

    g = dict(np.load(GOLD / f"{fixture}.npz"))
    net = networks.robertson() if name == "robertson" else networks.network_n40()
    model = sim.Model(net)
    sol = sim.Solver(**BENCH, ckpt_steps=ckpt)  # ckpt: adjoint checkpoints every 32 steps
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    args = (g["theta"], g["kfix"], g["ts"], g["meas"], g["sigma"])
    kw = dict(want_sx=(meth == 1))
    eager = _np(sim.simulate_batch(model, sol, *args, **kw))
    torch.cuda.synchronize()
    graph = sim.BatchGraph(model, sol, *args, **kw)
    try:
        for rep in range(3):
            out = _np(graph.launch())
            torch.cuda.synchronize()
            for k, v in eager.items():
                assert np.array_equal(out[k], v, equal_nan=True), (rep, k)
            # eager work of the program between replays
            sim.simulate_batch(model, sol, *args, **kw)
            torch.cuda.synchronize()
        # new parameters written into the graph's input buffer
        theta2 = g["theta"] + 0.05
        graph.set_inputs(theta=theta2)
        out = _np(graph.launch())
        ref = _np(sim.simulate_batch(model, sol, theta2, *args[1:], **kw))
        torch.cuda.synchronize()
        for k, v in ref.items():
            assert np.array_equal(out[k], v, equal_nan=True), ("theta2", k)
    finally:
        graph.close()
