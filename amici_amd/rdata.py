"""ReturnData outputs derived from simulated observables and their
sensitivities: residuals ``res``, residual sensitivities ``sres``, ``chi2``,
the Fisher information approximation ``FIM`` and the reporting modes that
select them.

Follows ReturnData::fres / fchi2 / fsres / fFIM (src/rdata.cpp:985-1204), the
failure invalidation ReturnData::invalidate (src/rdata.cpp:683-735), the
log10 chain rule of sres and FIM (src/rdata.cpp:826-837) and the field
selection per RDataReporting mode (src/rdata.cpp:85-160).  Noise
standard deviations come from the measurements' data, so their parameter
sensitivities ``ssigmay`` are zero; sensitivities are with respect to the
log10-scaled free parameters (the scale the kernels report sx in), which
makes the chain rule part of ``sy``.

All functions are vectorised over the batch and run on whatever device the
tensors live on (the simulation's device in ``run_simulations``).
"""
from __future__ import annotations

import enum

import math

import torch

from .networks import noise_scale


class RDataReporting(enum.IntEnum):
    """amici::RDataReporting (include/amici/defines.h:241-246)."""
    full = 0
    residuals = 1
    likelihood = 2
    observables_likelihood = 3


def reported_fields(mode: RDataReporting, forward: bool, sensi: bool) -> set:
    """Fields a ReturnData carries in ``mode`` (initialize_*_reporting,
    src/rdata.cpp:85-160) for a model with quadratic log-likelihood;
    ``forward`` = forward sensitivities, ``sensi`` = first order requested."""
    mode = RDataReporting(mode)
    out = set()
    fsa = forward and sensi
    if mode in (RDataReporting.full, RDataReporting.likelihood,
                RDataReporting.observables_likelihood):
        out |= {"llh"}
        if sensi:
            out |= {"sllh"}
        if fsa:
            out |= {"FIM"}
    if mode in (RDataReporting.full, RDataReporting.residuals,
                RDataReporting.observables_likelihood):
        out |= {"y", "sigmay"}
        if fsa:
            out |= {"sy", "ssigmay"}
    if mode in (RDataReporting.full, RDataReporting.residuals):
        out |= {"res", "chi2"}
        if fsa:
            out |= {"sres"}
    if mode == RDataReporting.full:
        out |= {"x"}
        if fsa:
            out |= {"sx"}
    return out


def residuals(y: torch.Tensor, meas: torch.Tensor, sigma: torch.Tensor, noise: int) -> torch.Tensor:
    """[nb, nt, ny]: (m - y) / sigma on the noise model's observable scale
    (linear, natural log or log10; fres, src/rdata.cpp:985-996); 0 where the
    measurement is missing."""
    obs = ~torch.isnan(meas)
    sc = noise_scale(noise)
    if sc == 1:
        r = (torch.log(meas) - torch.log(y)) / sigma
    elif sc == 2:
        r = (torch.log10(meas) - torch.log10(y)) / sigma
    else:
        r = (meas - y) / sigma
    return torch.where(obs, r, torch.zeros_like(r))


def residual_sensitivities(y: torch.Tensor, sy: torch.Tensor, meas: torch.Tensor,
                           sigma: torch.Tensor, noise: int) -> torch.Tensor:
    """[nb, nt, ny, np] from sy [nb, nt, np, ny]: -sy / sigma, -sy / (y sigma)
    on the natural-log scale, -sy / (y ln 10 sigma) on the log10 scale (fsres
    with ssigma = 0, src/rdata.cpp:1045-1062)."""
    syt = sy.transpose(-1, -2)                       # [nb, nt, ny, np]
    scale = 1.0 / sigma
    sc = noise_scale(noise)
    if sc == 1:
        scale = scale / y
    elif sc == 2:
        scale = scale / (y * math.log(10.0))
    s = -syt * scale.unsqueeze(-1)
    obs = ~torch.isnan(meas)
    return torch.where(obs.unsqueeze(-1), s, torch.zeros_like(s))


def sigma_residuals(sigma: torch.Tensor, meas: torch.Tensor, offset: float) -> torch.Tensor:
    """Error residuals sqrt(2 log sigma + offset) of measured entries
    (fres_error, src/rdata.cpp:998-1000), else 0."""
    r = torch.sqrt(2.0 * torch.log(sigma) + offset)
    return torch.where(~torch.isnan(meas), r, torch.zeros_like(r))


def first_invalid_time(y: torch.Tensor, status: torch.Tensor) -> torch.Tensor:
    """[nb] index of the first output time a failed simulation did not reach
    (its outputs are NaN from there on); nt for successful instances."""
    nb, nt = y.shape[:2]
    bad = torch.isnan(y).any(-1)                                  # [nb, nt]
    idx = torch.arange(nt, device=y.device).expand(nb, nt)
    cand = torch.cat([torch.where(bad, idx, nt), torch.full((nb, 1), nt, device=y.device)], 1)
    first = cand.amin(1)
    return torch.where(status != 0, first, torch.full_like(first, nt))


def derived_outputs(y: torch.Tensor, sy: torch.Tensor | None, meas: torch.Tensor,
                    sigma: torch.Tensor, status: torch.Tensor, noise: int, *,
                    sigma_res: bool = False, sigma_offset: float = 50.0,
                    fields: set | None = None) -> dict:
    """res [nb, k*nt*ny], sres [nb, k*nt*ny, np] (k = 2 with error residuals),
    chi2 [nb], FIM [nb, np, np] and ssigmay [nb, nt, np, ny], in ReturnData's
    layout, with failed instances invalidated as ReturnData::invalidate does:
    res / sres NaN from the first time point not reached, chi2 and FIM NaN
    when the forward problem failed (not after a backward-only failure)."""
    fields = fields if fields is not None else {"res", "sres", "chi2", "FIM", "ssigmay"}
    nb, nt, ny = y.shape
    out = {}
    it0 = first_invalid_time(y, status)                              # [nb]
    tmask = torch.arange(nt, device=y.device)[None, :] >= it0[:, None]  # [nb, nt] invalid
    # a failure of the forward problem (or of the pre-equilibration) leaves
    # output times unreached; a failure of the backward problem only
    # invalidates sllh (ReturnData::process_simulation_objects,
    # src/rdata.cpp:251-254), so llh, chi2 and the forward outputs stay
    failed = (status != 0) & (it0 < nt)
    nanf = torch.tensor(float("nan"), dtype=y.dtype, device=y.device)

    res = residuals(y, meas, sigma, noise)
    res = torch.where(tmask[..., None], nanf, res)
    blocks = [res.reshape(nb, nt * ny)]
    rerr = None
    if sigma_res:
        rerr = sigma_residuals(sigma, meas, sigma_offset)
        rerr = torch.where(tmask[..., None], nanf, rerr)
        blocks.append(rerr.reshape(nb, nt * ny))
    if "res" in fields:
        out["res"] = torch.cat(blocks, 1)
    if "chi2" in fields:
        c = (res * res).flatten(1).sum(1)
        if sigma_res:
            obs = ~torch.isnan(meas)
            c = c + torch.where(obs, rerr * rerr - sigma_offset, torch.zeros_like(rerr)).flatten(1).sum(1)
        out["chi2"] = torch.where(failed, nanf, c)
    if sy is not None:
        npar = sy.shape[2]
        if "ssigmay" in fields:
            out["ssigmay"] = torch.zeros_like(sy)
        sres = residual_sensitivities(y, sy, meas, sigma, noise)     # [nb, nt, ny, np]
        if "FIM" in fields:
            flat = sres.reshape(nb, nt * ny, npar)
            fim = flat.transpose(1, 2) @ flat
            out["FIM"] = torch.where(failed[:, None, None], nanf, fim)
        if "sres" in fields:
            sres = torch.where(tmask[..., None, None], nanf, sres)
            sb = [sres.reshape(nb, nt * ny, npar)]
            if sigma_res:
                # error residuals do not depend on the parameters (ssigma = 0)
                serr = torch.zeros_like(sres)
                sb.append(torch.where(tmask[..., None, None], nanf, serr).reshape(nb, nt * ny, npar))
            out["sres"] = torch.cat(sb, 1)
    return out
