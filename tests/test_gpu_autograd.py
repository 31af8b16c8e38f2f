"""GPU: the autograd boundary NLL(theta) against central finite differences
(the gradient_check recipe of python/sdist/amici/sim/sundials/gradient_check.py:
106-129: central differences on the log10 scale) and against the oracle's
forward-sensitivity gradient."""
from __future__ import annotations

import numpy as np
import pytest

from amici_amd import networks

pytestmark = pytest.mark.gpu
GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"
TIGHT = dict(rtol=1e-10, atol=1e-14, quad_rtol=1e-10, quad_atol=1e-14, maxsteps=200000,
             maxtime=60.0)


@pytest.mark.parametrize("graph", [False, True])
@pytest.mark.parametrize("fixture,netf", [("robertson_batch64", networks.robertson),
                                          ("n40_conditions16", networks.network_n40)])
def test_nll_gradient(gpu, fixture, netf, graph):
    import torch

    from amici_amd import autograd, sim

    g = dict(np.load(GOLD / f"{fixture}.npz"))
    nb = 4
    model = sim.Model(netf())
    dev = torch.device("cuda", 0)
    args = [torch.as_tensor(g[k][:nb], device=dev) for k in ("kfix",)] + \
           [torch.as_tensor(g["ts"], device=dev)] + \
           [torch.as_tensor(g[k][:nb], device=dev) for k in ("meas", "sigma")]
    solver = autograd.adjoint_solver(**TIGHT)
    theta = torch.as_tensor(g["theta"][:nb], device=dev).clone().requires_grad_(True)
    f = autograd.nll(theta, model, *args, solver=solver, graph=graph)
    assert torch.isfinite(f).all()
    f.sum().backward()
    grad = theta.grad.cpu().numpy()
    # NLL = -llh and its gradient = -sllh of the oracle
    np.testing.assert_allclose(f.detach().cpu().numpy(), -g["llh"][:nb], rtol=1e-6)
    scale = np.abs(g["sllh_fsa"][:nb]).max(1, keepdims=True)
    np.testing.assert_allclose(grad, -g["sllh_fsa"][:nb], rtol=1e-5, atol=1e-7 * scale.max())
    # central finite differences on the log10 scale for a few parameters
    eps = 1e-4
    js = np.argsort(-np.abs(grad[0]))[:3]
    for j in js:
        tp = theta.detach().clone()
        tm = theta.detach().clone()
        tp[:, j] += eps / 2
        tm[:, j] -= eps / 2
        fd = ((autograd.nll(tp, model, *args, solver=solver, graph=graph)
               - autograd.nll(tm, model, *args, solver=solver, graph=graph)) / eps).cpu().numpy()
        np.testing.assert_allclose(grad[:, j], fd, rtol=1e-3, atol=1e-6 * scale.max())
