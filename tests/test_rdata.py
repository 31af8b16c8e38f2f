"""ReturnData's derived outputs (amici_amd/rdata.py) against what the
likelihood fixes: res, sres, chi2 and FIM of ReturnData::fres / fchi2 /
fsres / fFIM (src/rdata.cpp:985-1204) are checked through torch.distributions'
Gaussian and log-normal log-densities, not through the formulas themselves:

* Gauss-Newton gradient: d llh / d theta = -sum res * sres when y is linear
  in theta (sigma does not depend on theta);
* FIM = Hessian of -llh for normal noise with y linear in theta, and for
  log-normal noise at zero residual;
* chi2 differences = -2 x llh differences; with error residuals
  (sqrt(2 log sigma + c)) -llh = chi2 / 2 + n log(2 pi) / 2 exactly;
* invalidation of failed instances (src/rdata.cpp:683-735).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from amici_amd import networks, rdata

NB, NT, NY, NP = 3, 5, 2, 4


def _case(noise, seed=0, missing=True):
    g = torch.Generator().manual_seed(seed)
    y0 = 0.5 + torch.rand(NB, NT, NY, generator=g, dtype=torch.float64)
    sy = torch.randn(NB, NT, NP, NY, generator=g, dtype=torch.float64)
    sigma = 0.05 + 0.2 * torch.rand(NB, NT, NY, generator=g, dtype=torch.float64)
    if noise == networks.NOISE_LOGNORMAL:
        meas = y0 * torch.exp(0.1 * torch.randn(NB, NT, NY, generator=g, dtype=torch.float64))
    else:
        meas = y0 + 0.1 * torch.randn(NB, NT, NY, generator=g, dtype=torch.float64)
    if missing:
        meas[0, 1, 0] = float("nan")
        meas[2, :, 1] = float("nan")
    return y0, sy, meas, sigma


def _llh(y, meas, sigma, noise):
    """Per-instance log-likelihood from torch.distributions over measured entries."""
    obs = ~torch.isnan(meas)
    m = torch.where(obs, meas, torch.ones_like(meas))
    if noise == networks.NOISE_LOGNORMAL:
        lp = torch.distributions.LogNormal(torch.log(y), sigma).log_prob(m)
    else:
        lp = torch.distributions.Normal(y, sigma).log_prob(m)
    return torch.where(obs, lp, torch.zeros_like(lp)).flatten(1).sum(1)


def _y_of(y0, sy, d):
    # y(theta0 + d) for y linear in theta: y0 + sum_p sy[.., p, :] d[b, p]
    return y0 + torch.einsum("btpy,bp->bty", sy, d)


NOISES = [networks.NOISE_NORMAL, networks.NOISE_LOGNORMAL]
OK = torch.zeros(NB, dtype=torch.int32)


@pytest.mark.parametrize("noise", NOISES)
def test_gauss_newton_gradient(noise):
    y0, sy, meas, sigma = _case(noise)
    d = torch.zeros(NB, NP, dtype=torch.float64, requires_grad=True)
    _llh(_y_of(y0, sy, d), meas, sigma, noise).sum().backward()
    out = rdata.derived_outputs(y0, sy, meas, sigma, OK, noise)
    grad = -torch.einsum("bk,bkp->bp", out["res"], out["sres"])
    torch.testing.assert_close(grad, d.grad, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("noise", NOISES)
def test_fim_is_the_hessian(noise):
    y0, sy, meas, sigma = _case(noise, seed=1)
    if noise == networks.NOISE_LOGNORMAL:
        meas = torch.where(torch.isnan(meas), meas, y0)      # zero residual: FIM exact
    out = rdata.derived_outputs(y0, sy, meas, sigma, OK, noise)
    for b in range(NB):
        def negllh(d, b=b):
            return -_llh(_y_of(y0[b:b + 1], sy[b:b + 1], d[None]), meas[b:b + 1],
                         sigma[b:b + 1], noise).sum()
        h = torch.autograd.functional.hessian(negllh, torch.zeros(NP, dtype=torch.float64))
        torch.testing.assert_close(out["FIM"][b], h, rtol=1e-10, atol=1e-10)
    fim = out["FIM"]
    torch.testing.assert_close(fim, fim.transpose(1, 2))
    assert torch.linalg.eigvalsh(fim).min() > -1e-10


@pytest.mark.parametrize("noise", NOISES)
def test_chi2_tracks_llh(noise):
    y0, sy, meas, sigma = _case(noise, seed=2)
    y1 = y0 * 1.03
    c0 = rdata.derived_outputs(y0, None, meas, sigma, OK, noise)["chi2"]
    c1 = rdata.derived_outputs(y1, None, meas, sigma, OK, noise)["chi2"]
    dl = _llh(y1, meas, sigma, noise) - _llh(y0, meas, sigma, noise)
    torch.testing.assert_close(c1 - c0, -2.0 * dl, rtol=1e-12, atol=1e-12)


def test_sigma_residuals_complete_the_likelihood():
    noise = networks.NOISE_NORMAL
    y0, sy, meas, sigma = _case(noise, seed=3)
    sigma = sigma + 1.0          # log sigma > 0 keeps sqrt(2 log sigma + 0) real
    for offset in (0.0, 50.0):
        out = rdata.derived_outputs(y0, sy, meas, sigma, OK, noise, sigma_res=True,
                                    sigma_offset=offset)
        n = (~torch.isnan(meas)).flatten(1).sum(1).to(torch.float64)
        torch.testing.assert_close(-_llh(y0, meas, sigma, noise),
                                   0.5 * out["chi2"] + 0.5 * n * math.log(2 * math.pi),
                                   rtol=1e-12, atol=1e-12)
        assert out["res"].shape == (NB, 2 * NT * NY)
        assert out["sres"].shape == (NB, 2 * NT * NY, NP)
        assert torch.all(out["sres"][:, NT * NY:] == 0)
        plain = rdata.derived_outputs(y0, sy, meas, sigma, OK, noise)
        torch.testing.assert_close(out["res"][:, :NT * NY], plain["res"])
        torch.testing.assert_close(out["FIM"], plain["FIM"])


def test_missing_measurements_give_zero_residuals():
    y0, sy, meas, sigma = _case(networks.NOISE_NORMAL)
    out = rdata.derived_outputs(y0, sy, meas, sigma, OK, networks.NOISE_NORMAL)
    res = out["res"].reshape(NB, NT, NY)
    sres = out["sres"].reshape(NB, NT, NY, NP)
    miss = torch.isnan(meas)
    assert torch.all(res[miss] == 0) and torch.all(sres[miss] == 0)
    assert torch.all(res[~miss] != 0)
    assert torch.all(out["ssigmay"] == 0) and out["ssigmay"].shape == sy.shape


def test_failed_instances_are_invalidated():
    noise = networks.NOISE_NORMAL
    y0, sy, meas, sigma = _case(noise, missing=False)
    status = torch.tensor([0, -3, -1], dtype=torch.int32)
    y = y0.clone()
    sy = sy.clone()
    y[1, 2:] = float("nan")          # the kernels write NaN from the first time point not reached
    sy[1, 2:] = float("nan")
    y[2] = float("nan")
    sy[2] = float("nan")
    out = rdata.derived_outputs(y, sy, meas, sigma, status, noise, sigma_res=True)
    assert rdata.first_invalid_time(y, status).tolist() == [NT, 2, 0]
    res = out["res"].reshape(NB, 2, NT, NY)
    sres = out["sres"].reshape(NB, 2, NT, NY, NP)
    assert torch.isfinite(res[0]).all() and torch.isfinite(sres[0]).all()
    assert torch.isfinite(res[1, :, :2]).all() and torch.isnan(res[1, :, 2:]).all()
    assert torch.isfinite(sres[1, :, :2]).all() and torch.isnan(sres[1, :, 2:]).all()
    assert torch.isnan(res[2]).all() and torch.isnan(sres[2]).all()
    assert math.isfinite(out["chi2"][0]) and torch.isnan(out["chi2"][1:]).all()
    assert torch.isfinite(out["FIM"][0]).all() and torch.isnan(out["FIM"][1:]).all()


def test_backward_failure_keeps_forward_outputs():
    """A failure of the backward problem only (status set, every output time
    reached) invalidates sllh alone in the reference (src/rdata.cpp:251-254):
    residuals, chi2 and FIM of the forward outputs stay."""
    noise = networks.NOISE_NORMAL
    y0, sy, meas, sigma = _case(noise, missing=False)
    ok = rdata.derived_outputs(y0, sy, meas, sigma, OK, noise)
    status = torch.tensor([0, -3, 0], dtype=torch.int32)
    out = rdata.derived_outputs(y0, sy, meas, sigma, status, noise)
    assert rdata.first_invalid_time(y0, status).tolist() == [NT, NT, NT]
    for k in ("res", "sres", "chi2", "FIM"):
        assert torch.equal(out[k], ok[k]), k


def test_reporting_modes():
    F = rdata.RDataReporting
    full = rdata.reported_fields(F.full, forward=True, sensi=True)
    assert full >= {"x", "y", "sx", "sy", "llh", "sllh", "res", "sres", "chi2", "FIM"}
    assert rdata.reported_fields(F.full, forward=False, sensi=True) & {"sx", "sy", "sres", "FIM"} == set()
    resid = rdata.reported_fields(F.residuals, forward=True, sensi=True)
    assert "llh" not in resid and "x" not in resid and {"res", "sres", "sy"} <= resid
    lik = rdata.reported_fields(F.likelihood, forward=True, sensi=True)
    assert lik == {"llh", "sllh", "FIM"}
    ol = rdata.reported_fields(F.observables_likelihood, forward=True, sensi=True)
    assert ol == {"llh", "sllh", "FIM", "y", "sigmay", "sy", "ssigmay"}
    assert rdata.reported_fields(F.likelihood, forward=False, sensi=False) == {"llh"}


def test_fields_subset_is_respected():
    y0, sy, meas, sigma = _case(networks.NOISE_NORMAL)
    out = rdata.derived_outputs(y0, sy, meas, sigma, OK, networks.NOISE_NORMAL, fields={"FIM"})
    assert set(out) == {"FIM"}
    np.testing.assert_array_equal(out["FIM"].shape, (NB, NP, NP))
