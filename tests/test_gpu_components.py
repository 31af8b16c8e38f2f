"""GPU unit tests of the integrator's building blocks (team reductions and
broadcasts, row-per-lane LU, mass-action model evaluation) against numpy and
the oracle's model functions, for a 64-lane team with 3 states (Robertson), a
64-lane team with 40 states (40-species network) and a 256-lane workgroup team
with 200 states (200-species network: cross-wavefront reductions through LDS,
blocked LU in global memory)."""
import ctypes

import numpy as np
import pytest
import torch

import component_build
from amici_amd import networks
from oracle import cvodes_oracle as O

NETS = {"robertson": networks.robertson, "n40": networks.network_n40,
        "n200": networks.network_n200}


def _lib(name):
    net = NETS[name]()
    lib = ctypes.CDLL(str(component_build.build(net)))
    d = (ctypes.c_int * 4)()
    lib.amxt_dims(d)
    return net, lib, list(d)


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(NETS))
def test_team_ops(gpu, name):
    net, lib, (nx, nr, npar, T) = _lib(name)
    nteams = 37
    rng = np.random.default_rng(1)
    v = rng.standard_normal((nteams, T))
    v[3, :] = 1.5  # ties: argmax must pick the first lane
    vin = torch.tensor(v, device=gpu)
    out = torch.zeros((nteams, T, 6), dtype=torch.float64, device=gpu)
    assert lib.amxt_team(_p(vin), _p(out), nteams, None) == 0
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    np.testing.assert_allclose(o[:, :, 0], np.repeat(v.sum(1, keepdims=True), T, 1), rtol=1e-13)
    np.testing.assert_array_equal(o[:, :, 1], np.repeat(v.max(1, keepdims=True), T, 1))
    np.testing.assert_array_equal(o[:, :, 3], np.repeat(v.argmax(1)[:, None], T, 1))
    np.testing.assert_array_equal(o[:, :, 4], np.repeat(v[:, 1 % T][:, None], T, 1))
    np.testing.assert_array_equal(o[:, :, 5], v[:, ::-1])
    # every lane of a team holds the bitwise-same reduction
    assert np.all(o[:, :, 0] == o[:, :1, 0])


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(NETS))
def test_lu_solve(gpu, name):
    net, lib, (nx, nr, npar, T) = _lib(name)
    nteams = 53
    rng = np.random.default_rng(2)
    A = rng.standard_normal((nteams, nx, nx)) + 0.5 * np.eye(nx)
    A[0] = np.eye(nx)[::-1] * 2.0 + 1e-3  # pivoting required
    b = rng.standard_normal((nteams, nx))
    At, bt = torch.tensor(A, device=gpu), torch.tensor(b, device=gpu)
    x = torch.zeros((nteams, nx), dtype=torch.float64, device=gpu)
    ok = torch.zeros(nteams, dtype=torch.int32, device=gpu)
    work = torch.zeros((nteams, nx, nx), dtype=torch.float64, device=gpu)
    assert lib.amxt_lu(_p(At), _p(bt), _p(x), _p(ok), _p(work), nteams, None) == 0
    torch.cuda.synchronize()
    assert ok.cpu().numpy().all()
    xr = np.linalg.solve(A, b[..., None])[..., 0]
    cond = np.linalg.cond(A)
    err = np.abs(x.cpu().numpy() - xr).max(1) / np.abs(xr).max(1)
    assert np.all(err < 1e-13 * cond), (err / cond).max()


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(NETS))
def test_model_eval(gpu, name):
    net, lib, (nx, nr, npar, T) = _lib(name)
    nteams = 29
    rng = np.random.default_rng(3)
    theta = net.theta_nominal[None, :] + 0.3 * rng.standard_normal((nteams, npar))
    x = rng.uniform(0.0, 2.0, (nteams, nx))
    v = rng.standard_normal((nteams, nx))
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), device=gpu)
    tt, xt, vt = dev(theta), dev(x), dev(v)
    xdot = torch.zeros((nteams, nx), dtype=torch.float64, device=gpu)
    J = torch.zeros((nteams, nx, nx), dtype=torch.float64, device=gpu)
    u = torch.zeros((nteams, nr), dtype=torch.float64, device=gpu)
    jtv = torch.zeros((nteams, nx), dtype=torch.float64, device=gpu)
    assert lib.amxt_model(_p(tt), _p(xt), _p(vt), _p(xdot), _p(J), _p(u), _p(jtv), nteams, None) == 0
    torch.cuda.synchronize()
    om = O.OracleModel(net)
    S = net.stoichiometry()
    for k in range(nteams):
        fr = om.xdot(theta[k], x[k])
        Jr = om.jac(theta[k], x[k])
        scale = np.abs(Jr).max() * np.abs(x[k]).max() + 1e-300
        np.testing.assert_allclose(xdot.cpu().numpy()[k], fr, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(J.cpu().numpy()[k], Jr, rtol=1e-13, atol=1e-13 * np.abs(Jr).max())
        np.testing.assert_allclose(u.cpu().numpy()[k], S.T @ v[k], rtol=1e-13, atol=1e-13)
        np.testing.assert_allclose(jtv.cpu().numpy()[k], Jr.T @ v[k], rtol=1e-12,
                                   atol=1e-13 * np.abs(Jr).max() * np.abs(v[k]).max())
