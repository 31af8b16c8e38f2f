"""The C-ABI model modules load on the CPU and export every entry point that
include/amx.h declares; host-side argument checks and workspace sizing work
without touching the GPU."""
from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest

from amici_amd import _lib, codegen, networks

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "amx.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*)\s+(amx_\w+)\s*\(", text, re.M)))


@pytest.fixture(scope="module", params=["robertson", "n40", "linear3", "n40open", "n200", "n200open"])
def lib(request):
    net = {"robertson": networks.robertson, "n40": networks.network_n40,
           "linear3": lambda: networks.linear_chain(3),
           "n40open": networks.network_n40_open, "n200": networks.network_n200,
           "n200open": networks.network_n200_open}[request.param]()
    so = codegen.compile_model(net)
    return net, _lib.ModelLib(so)


def test_header_declares_the_binding_symbols():
    assert set(_declared()) == set(_lib.SYMBOLS)


def test_exports_and_dims(lib):
    net, L = lib
    for s in _declared():
        assert hasattr(L.lib, s), s
    assert (L.nx, L.nr, L.np, L.nk, L.ny) == (net.nx, net.nr, net.np, net.nk, net.ny)
    assert L.team == (64 if net.nx <= 64 else 256) and L.noise == int(net.noise)
    assert L.status_string(-1) == "AMICI_TOO_MUCH_WORK"
    assert L.status_string(-1001) == "AMICI_NOT_RUN"


def test_workspace_bytes(lib):
    net, L = lib
    o = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, maxsteps=1000)
    b1 = L.workspace_bytes(o, 1, 10)
    b8 = L.workspace_bytes(o, 8, 10)
    # trajectory: t, y and the order per stored step and instance dominate
    # (polynomial interpolation, the default)
    assert b8 - b1 >= 7 * 1001 * ((1 + net.nx) * 8 + 4)
    o.max_traj = 100
    assert L.workspace_bytes(o, 8, 10) < b8
    f = _lib.make_opts(sensi_meth=_lib.SENSI_FORWARD)
    assert L.workspace_bytes(f, 4, 10) >= 4 * net.np * net.nx * 8
    # states only: counters, plus the Newton matrices of a workgroup-team model
    lu = 4 * net.nx * net.nx * 8 if net.nx > 64 else 0
    assert lu <= L.workspace_bytes(_lib.make_opts(), 4, 10) <= 512 + lu


def test_invalid_arguments_rejected_on_host(lib):
    net, L = lib
    fn = L.lib.amx_run_batch
    o = _lib.make_opts(sensi_meth=_lib.SENSI_FORWARD)
    ins = _lib.BatchIn(1, 3, 0, None, None, None, None, None, None)
    outs = _lib.BatchOut()
    assert fn(ctypes.byref(o), ctypes.byref(ins), ctypes.byref(outs), None, 0, None) == -22
    assert fn(None, ctypes.byref(ins), ctypes.byref(outs), None, 0, None) == -22
    empty = _lib.BatchIn(0, 3, 0, None, None, None, None, None, None)
    assert fn(ctypes.byref(o), ctypes.byref(empty), ctypes.byref(outs), None, 0, None) == 0
    # a non-null but too-small workspace is refused before any launch
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None)
    st = _lib.make_opts()
    assert fn(ctypes.byref(st), ctypes.byref(ins), ctypes.byref(outs), p, 8, None) == -23
    bad = _lib.make_opts(rtol=-1.0)
    assert fn(ctypes.byref(bad), ctypes.byref(ins), ctypes.byref(outs), p, 8, None) == -22
    bad = _lib.make_opts(rtol=math.nan)
    assert fn(ctypes.byref(bad), ctypes.byref(ins), ctypes.byref(outs), p, 8, None) == -22
    # pre-equilibration needs fixed parameters and a non-negative Newton budget
    pre = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, p)
    if net.nk == 0:
        assert fn(ctypes.byref(o), ctypes.byref(pre), ctypes.byref(outs), p, 1 << 30, None) == -22
    neg = _lib.make_opts(sensi_meth=_lib.SENSI_FORWARD, newton_maxsteps=-1)
    if net.nk:
        assert fn(ctypes.byref(neg), ctypes.byref(pre), ctypes.byref(outs), p, 1 << 30, None) == -22
    # the sequential path has no adjoint
    adj = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT)
    assert L.lib.amx_run_sequential(ctypes.byref(adj), ctypes.byref(ins), ctypes.byref(outs),
                                    p, 1 << 30, None) == -22


def test_timing_read_without_launches(lib):
    _, L = lib
    L.set_timing(True)
    t = L.kernel_times()
    L.set_timing(False)
    assert t["fwd_launches"] == 0 and t["bwd_launches"] == 0 and t["dropped"] == 0


def test_tables_match_sympy():
    """The constexpr sparsity tables emitted into the model module, read the
    way the device code reads them, agree with the sympy derivation of xdot,
    J, J^T v, dxdot/dp and y."""
    rng = np.random.default_rng(1)
    for net in (networks.robertson(), networks.network_n40(), networks.linear_chain(3)):
        t = codegen.model_tables(net)
        sm = codegen.symbolic_model(net)
        x = rng.uniform(0.1, 2.0, net.nx)
        k = 10.0 ** net.theta_nominal
        v = rng.standard_normal(net.nx)
        subs = {**dict(zip(sm["x"], x)), **dict(zip(sm["p"], k))}
        xv = lambda i: 1.0 if i < 0 else x[i]  # noqa: E731
        w = np.array([k[t["KI"][r]] * xv(t["RA"][r]) * xv(t["RB"][r]) for r in range(net.nr)])
        f = np.array([sum(t["SP_C"][e] * w[t["SP_R"][e]]
                          for e in range(t["SP_PTR"][i], t["SP_PTR"][i + 1]))
                      for i in range(net.nx)])
        f_sym = np.array([float(e.subs(subs)) for e in sm["xdot"]])
        np.testing.assert_allclose(f, f_sym, rtol=1e-12, atol=1e-12 * np.abs(f_sym).max())
        J = np.zeros((net.nx, net.nx))
        for q in range(t["JNZ"]):
            for e in range(t["JT_PTR"][q], t["JT_PTR"][q + 1]):
                J[t["J_I"][q], t["J_J"][q]] += t["JT_C"][e] * k[t["KI"][t["JT_R"][e]]] * xv(t["JT_O"][e])
        J_sym = np.array(sm["J"].subs(subs).evalf(), dtype=float)
        np.testing.assert_allclose(J, J_sym, rtol=1e-12, atol=1e-12 * np.abs(J_sym).max())
        # J^T v through the per-state reaction list (dw_r/dx_s) and S columns
        jtv = np.zeros(net.nx)
        for s_ in range(net.nx):
            for e in range(t["DX_PTR"][s_], t["DX_PTR"][s_ + 1]):
                r = t["DX_R"][e]
                sv = sum(t["RS_C"][q] * v[t["RS_S"][q]] for q in range(t["RS_PTR"][r], t["RS_PTR"][r + 1]))
                jtv[s_] += t["DX_C"][e] * k[t["KI"][r]] * xv(t["DX_O"][e]) * sv
        np.testing.assert_allclose(jtv, J_sym.T @ v, rtol=1e-10, atol=1e-12 * np.abs(J_sym).max())
        # dxdot/dp_j = sum over the reactions of rate j of S[:, r] w_r / k_j
        dfdp_sym = np.array(sm["dfdp"].subs(subs).evalf(), dtype=float)
        S = net.stoichiometry()
        for j in range(net.np):
            col = sum(S[:, r] * w[r] / k[j] for r in t["PR_R"][t["PR_PTR"][j]:t["PR_PTR"][j + 1]])
            np.testing.assert_allclose(col, dfdp_sym[:, j], rtol=1e-12, atol=1e-14)
        y = np.array([sum(t["OB_C"][e] * x[t["OB_S"][e]] for e in range(t["OB_PTR"][q], t["OB_PTR"][q + 1]))
                      for q in range(net.ny)])
        y_sym = np.array([float(e.subs(subs)) for e in sm["y"]])
        np.testing.assert_allclose(y, y_sym, rtol=1e-14)


def test_teams_narrower_than_a_wavefront_are_refused():
    """A team is one 64-lane wavefront or a workgroup of 128 / 256 lanes;
    narrower teams (several instances per wavefront) are not provided."""
    net = networks.robertson()
    for team in (1, 4, 32):
        with pytest.raises(ValueError, match="team size"):
            codegen.emit_model_traits(net, team)
    with pytest.raises(ValueError, match="team size"):
        codegen.emit_model_traits(networks.network_n40(), 32)


def test_abi_v10_struct_layout():
    """amx_solver_opts ends with max_blocks, interp, stldet, failfast and
    ckpt_steps, amx_batch_in with the parameter list, scales, initial states,
    their sensitivities and the start times, amx_batch_out with diag; 18
    counters per instance (include/amx.h, ABI version 10)."""
    assert _lib.AMX_ABI_VERSION == 10 and _lib.AMX_NSTATS == 18
    assert [f[0] for f in _lib.BatchIn._fields_[-6:]] == ["nplist", "plist", "pscale", "x0", "sx0",
                                                          "t0"]
    assert [f[0] for f in _lib.SolverOpts._fields_[-5:]] == ["max_blocks", "interp", "stldet",
                                                             "failfast", "ckpt_steps"]
    assert _lib.make_opts().failfast == 0 and _lib.make_opts(failfast=True).failfast == 1
    assert _lib.make_opts().ckpt_steps == 0 and _lib.make_opts(ckpt_steps=32).ckpt_steps == 32
    assert _lib.BatchOut._fields_[-1][0] == "diag"
    text = (ROOT / "include" / "amx.h").read_text()
    assert "#define AMX_ABI_VERSION 10" in text and "#define AMX_NDIAG 9" in text
    assert "#define AMX_NSTATS 18" in text
    assert "#define AMX_INTERP_HERMITE 1" in text and "#define AMX_INTERP_POLYNOMIAL 2" in text
    assert "#define AMX_STLDET_ON 1" in text and "#define AMX_STLDET_OFF 2" in text
    assert _lib.make_opts().stldet == _lib.STLDET_ON
    assert _lib.make_opts(stldet=False).stldet == _lib.STLDET_OFF


def test_checkpoint_workspace(lib):
    """With a checkpoint every K forward steps the workspace holds one
    interval of K + 1 points, a record of 6 nx + 72 doubles per checkpoint and
    the backward problem's LDS block per instance, instead of every step."""
    net, L = lib
    nb, K = 8, 32
    full = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, maxsteps=1000)
    ck = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, maxsteps=1000, ckpt_steps=K)
    per_pt = (1 + net.nx) * 8 + 4
    rec = nb * (1000 // K + 1) * (6 * net.nx + 72) * 8
    d = L.workspace_bytes(full, nb, 10) - L.workspace_bytes(ck, nb, 10)
    spill = nb * (1001 - (K + 1)) * per_pt - rec - d
    assert 0 < spill <= nb * 65536 + 4096, spill
    # states and forward sensitivities ignore the setting
    for m in (_lib.SENSI_NONE, _lib.SENSI_FORWARD):
        assert (L.workspace_bytes(_lib.make_opts(sensi_meth=m, ckpt_steps=K), nb, 10)
                == L.workspace_bytes(_lib.make_opts(sensi_meth=m), nb, 10))
    bad = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, ckpt_steps=-1)
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    assert L.lib.amx_run_batch(ctypes.byref(bad), ctypes.byref(ins), ctypes.byref(outs), p, 8,
                               None) == -22


def test_solver_checkpoint_setting():
    from amici_amd import sim
    s = sim.Solver()
    assert s.get_checkpoint_steps() == 0 and s.c_opts().ckpt_steps == 0
    s.set_checkpoint_steps(32)
    assert s.c_opts().ckpt_steps == 32
    with pytest.raises(ValueError):
        s.set_checkpoint_steps(-1)


def test_invalid_stldet_rejected(lib):
    net, L = lib
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    o = _lib.make_opts()
    o.stldet = 3
    assert L.lib.amx_run_batch(ctypes.byref(o), ctypes.byref(ins), ctypes.byref(outs),
                               p, 8, None) == -22


def test_interpolation_workspace(lib):
    """Polynomial interpolation stores the order per step instead of y'."""
    net, L = lib
    poly = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, maxsteps=1000)
    herm = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, maxsteps=1000, interp=_lib.INTERP_HERMITE)
    d = L.workspace_bytes(herm, 8, 10) - L.workspace_bytes(poly, 8, 10)
    assert abs(d - 8 * 1001 * (8 * net.nx - 4)) <= 512
    bad = _lib.make_opts(sensi_meth=_lib.SENSI_ADJOINT, interp=3)
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    assert L.lib.amx_run_batch(ctypes.byref(bad), ctypes.byref(ins), ctypes.byref(outs), p, 8,
                               None) == -22


def test_negative_max_blocks_rejected(lib):
    net, L = lib
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    o = _lib.make_opts(max_blocks=-1)
    assert L.lib.amx_run_batch(ctypes.byref(o), ctypes.byref(ins), ctypes.byref(outs),
                               p, 8, None) == -22


def test_parameter_list_length_rejected(lib):
    """A parameter list longer than np is refused before any launch."""
    net, L = lib
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p if net.nk else None, p, None, p, None, net.np + 1, p, None,
                       None)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    o = _lib.make_opts()
    assert L.lib.amx_run_batch(ctypes.byref(o), ctypes.byref(ins), ctypes.byref(outs),
                               p, 8, None) == -22


def test_model_parameter_list_and_scale_api():
    """Model.set_parameter_list / set_parameter_scale / set_initial_state
    (include/amici/model.h:414-423, 954) validate their inputs; ExpData
    carries plist, pscale and x0 (empty = the model's)."""
    from amici_amd import sim
    m = sim.Model(networks.linear_chain(3))
    assert m.get_parameter_list() == list(range(m.np)) and m.nplist() == m.np
    m.set_parameter_list([2, 0])
    assert m.get_parameter_list() == [2, 0] and m.nplist() == 2
    with pytest.raises(ValueError):
        m.set_parameter_list([m.np])
    m.set_parameter_scale("ln")
    assert m.get_parameter_scale() == [_lib.PSCALE_LN] * m.np
    m.set_parameter_scale(["none", 1, "log10", 2])
    assert m.get_parameter_scale() == [0, 1, 2, 2]
    with pytest.raises(ValueError):
        m.set_parameter_scale([0, 1])
    with pytest.raises(ValueError):
        m.set_parameter_scale("log2")
    x0 = np.arange(m.nx_rdata, dtype=float)
    m.set_initial_state(x0)
    assert np.array_equal(m.get_initial_state(), x0)
    m.set_initial_state([])
    assert np.array_equal(m.get_initial_state(), np.asarray(m.network.x0_const, float))
    e = sim.ExpData(m)
    assert list(e.plist) == [] and list(e.pscale) == [] and len(e.x0) == 0


def test_sx0_with_preequilibration_rejected(lib):
    """Caller-given initial state sensitivities together with a
    pre-equilibration are refused before any launch."""
    net, L = lib
    if not net.nk:
        return
    buf = (ctypes.c_double * 4)()
    p = ctypes.cast(buf, ctypes.c_void_p).value
    ins = _lib.BatchIn(1, 3, 0, p, p, p, None, p, p, 0, None, None, None, p)
    outs = _lib.BatchOut(None, None, None, p, p, p, None, None)
    o = _lib.make_opts(sensi_meth=_lib.SENSI_FORWARD)
    assert L.lib.amx_run_batch(ctypes.byref(o), ctypes.byref(ins), ctypes.byref(outs),
                               p, 8, None) == -22


def test_model_start_time_setting():
    from amici_amd import sim
    m = sim.Model(networks.robertson())
    assert m.t0() == 0.0
    m.set_t0(2.5)
    assert m.t0() == 2.5
    e = sim.ExpData(m)
    assert e.t_start is None


def test_condition_start_times_and_presimulation():
    """run_simulations' start times and fixed parameters: ExpData t_start,
    else the model's t0; a pre-simulation starts t_presim earlier with the
    pre-simulation's fixed parameters (the condition's own when those are
    empty)."""
    from amici_amd import sim
    m = sim.Model(networks.network_n40())
    m.set_t0(1.5)
    k = np.arange(m.nk, dtype=float) + 1.0
    a, b, c, d = (sim.ExpData(m) for _ in range(4))
    a.fixed_parameters = k
    b.fixed_parameters, b.t_start = k, 4.0
    c.fixed_parameters, c.t_presim, c.fixed_parameters_presimulation = k, 2.0, 10 * k
    d.fixed_parameters, d.t_presim, d.t_start = k, 3.0, 0.0
    kfix, t0 = sim.condition_start(m, [a, b, c, d])
    np.testing.assert_array_equal(t0, [1.5, 4.0, -0.5, -3.0])
    np.testing.assert_array_equal(kfix, np.stack([k, k, 10 * k, k]))
