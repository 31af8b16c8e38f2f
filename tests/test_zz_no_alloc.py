"""GPU: amx_run_batch_static, the entry point that allocates nothing (its
workspace is a static device array of the library), against amx_run_batch
through the C ABI: two calls in a row, the second with a smaller batch, equal
bit for bit; a batch whose workspace exceeds the static array is refused."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np
import pytest

from amici_amd import _lib, networks

GOLD = Path(__file__).resolve().parent / "golden"
BENCH = dict(rtol=1e-8, atol=1e-12, quad_rtol=1e-8, quad_atol=1e-12, maxsteps=10000)

pytestmark = pytest.mark.gpu


def _structs(model, solver, g, nb, dev):
    import torch
    from amici_amd import sim
    p = sim._prepare(model, solver, g["theta"][:nb], g["kfix"][:nb], g["ts"], g["meas"][:nb],
                     g["sigma"][:nb], want_x=True, want_y=True, want_sx=False, want_stats=True,
                     kfix_preeq=None)
    return p


def _host(p):
    return {k: v.cpu().numpy() for k, v in p["out"].items() if v is not None}


@pytest.mark.parametrize("name,fixture,meth,max_traj", [
    ("robertson", "robertson_batch64", 2, 4000),
    ("robertson", "robertson_batch64", 1, 0),
    ("n40", "n40_conditions16", 2, 2500),
    ("n40", "n40_conditions16", 0, 0)])
def test_static_workspace_equals_allocating_entry_point(gpu, name, fixture, meth, max_traj):
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / f"{fixture}.npz"))
    net = networks.robertson() if name == "robertson" else networks.network_n40()
    model = sim.Model(net)
    L = model.lib.lib
    sol = sim.Solver(**BENCH, max_traj=max_traj)
    if meth:
        sol.set_sensitivity_order(sim.SensitivityOrder.first)
        sol.set_sensitivity_method(meth)
    nb_full = g["theta"].shape[0]
    stream = torch.cuda.current_stream()
    cap = L.amx_static_workspace_bytes()
    assert cap == 64 << 20
    for nb in (nb_full, nb_full // 4):
        # static entry point
        ps = _structs(model, sol, g, nb, gpu)
        need = model.lib.workspace_bytes(ps["opts"], nb, ps["nt"])
        assert need <= cap, (need, cap)
        rc = L.amx_run_batch_static(ctypes.byref(ps["opts"]), ctypes.byref(ps["bin"]),
                                    ctypes.byref(ps["bout"]), ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, rc
        # allocating entry point with its own workspace
        pa = _structs(model, sol, g, nb, gpu)
        ws = torch.empty(max(need, 256), dtype=torch.uint8, device=gpu)
        rc = L.amx_run_batch(ctypes.byref(pa["opts"]), ctypes.byref(pa["bin"]),
                             ctypes.byref(pa["bout"]), ws.data_ptr(), ws.numel(),
                             ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, rc
        torch.cuda.synchronize()
        hs, ha = _host(ps), _host(pa)
        assert (ha["status"] == 0).all(), ha["status"]
        for k in ha:
            assert np.array_equal(hs[k], ha[k], equal_nan=True), (nb, k)


def test_static_workspace_limit_refused(gpu):
    """An adjoint batch whose trajectory storage exceeds the static array is
    refused with -24 before anything is launched."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / "n40_conditions16.npz"))
    model = sim.Model(networks.network_n40())
    # max_traj = maxsteps + 1: 30001 stored steps of (t, x, order), 10 MB per instance
    sol = sim.Solver(**{**BENCH, "maxsteps": 30000})
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    p = _structs(model, sol, g, 16, gpu)
    assert model.lib.workspace_bytes(p["opts"], 16, p["nt"]) > model.lib.lib.amx_static_workspace_bytes()
    rc = model.lib.lib.amx_run_batch_static(ctypes.byref(p["opts"]), ctypes.byref(p["bin"]),
                                            ctypes.byref(p["bout"]),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -24
    torch.cuda.synchronize()
    assert (p["out"]["status"].cpu().numpy() == sim.AMICI_NOT_RUN).all()


def test_static_workspace_with_checkpoints(gpu):
    """With adjoint checkpoints every 32 steps the static array holds 64 N40
    conditions at maxsteps 10000 (storing every step they would need 212 MB);
    two calls in a row, 64 then 16 conditions, equal amx_run_batch bit for bit."""
    import torch
    from amici_amd import sim
    g = dict(np.load(GOLD / "n40_conditions16.npz"))
    g = {k: (np.concatenate([v] * 4) if k in ("theta", "kfix", "meas", "sigma") else v)
         for k, v in g.items()}
    model = sim.Model(networks.network_n40())
    L = model.lib.lib
    full = sim.Solver(**BENCH)
    full.set_sensitivity_order(sim.SensitivityOrder.first)
    full.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    sol = sim.Solver(**BENCH, ckpt_steps=32)
    sol.set_sensitivity_order(sim.SensitivityOrder.first)
    sol.set_sensitivity_method(sim.SensitivityMethod.adjoint)
    stream = torch.cuda.current_stream()
    cap = L.amx_static_workspace_bytes()
    pf = _structs(model, full, g, 64, gpu)
    assert model.lib.workspace_bytes(pf["opts"], 64, pf["nt"]) > cap
    for nb in (64, 16):
        ps = _structs(model, sol, g, nb, gpu)
        need = model.lib.workspace_bytes(ps["opts"], nb, ps["nt"])
        assert need <= cap, (need, cap)
        rc = L.amx_run_batch_static(ctypes.byref(ps["opts"]), ctypes.byref(ps["bin"]),
                                    ctypes.byref(ps["bout"]), ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, rc
        pa = _structs(model, sol, g, nb, gpu)
        ws = torch.empty(need, dtype=torch.uint8, device=gpu)
        rc = L.amx_run_batch(ctypes.byref(pa["opts"]), ctypes.byref(pa["bin"]),
                             ctypes.byref(pa["bout"]), ws.data_ptr(), ws.numel(),
                             ctypes.c_void_p(stream.cuda_stream))
        assert rc == 0, rc
        torch.cuda.synchronize()
        hs, ha = _host(ps), _host(pa)
        assert (ha["status"] == 0).all(), ha["status"]
        for k in ha:
            assert np.array_equal(hs[k], ha[k], equal_nan=True), (nb, k)
