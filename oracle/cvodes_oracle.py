"""TEST INFRASTRUCTURE ONLY — Python binding of the oracle library.

The oracle is the reference's own SUNDIALS CVODES 7.1.1 (compiled from
/root/reference/ThirdParty/sundials by oracle/Makefile into
oracle/_ref/libamx_oracle.so) driven by oracle/cvodes_harness.c, a C
restatement of AMICI's forward / backward problem, objective and chain-rule
layer.  Only tests, ``__graft_entry__.smoke()`` and the cpu_baseline leg of
``bench.py`` may use it, and only as the checker or the reported CPU baseline.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libamx_oracle.so"
NCOUNTERS = 16
CV_SIMULTANEOUS, CV_STAGGERED = 1, 2
CV_HERMITE, CV_POLYNOMIAL = 1, 2


class _Model(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("nr", ctypes.c_int), ("np", ctypes.c_int),
                ("nk", ctypes.c_int), ("ny", ctypes.c_int),
                ("rptr", ctypes.c_void_p), ("ridx", ctypes.c_void_p),
                ("pptr", ctypes.c_void_p), ("pidx", ctypes.c_void_p),
                ("kidx", ctypes.c_void_p), ("x0c", ctypes.c_void_p),
                ("x0k", ctypes.c_void_p), ("C", ctypes.c_void_p), ("noise", ctypes.c_int)]


class _Opts(ctypes.Structure):
    _fields_ = [("rtol", ctypes.c_double), ("atol", ctypes.c_double),
                ("quad_rtol", ctypes.c_double), ("quad_atol", ctypes.c_double),
                ("maxsteps", ctypes.c_long), ("sensi", ctypes.c_int), ("ism", ctypes.c_int),
                ("interp", ctypes.c_int), ("stldet", ctypes.c_int), ("t_preeq", ctypes.c_double),
                ("nplist", ctypes.c_int), ("plist", ctypes.c_void_p), ("pscale", ctypes.c_void_p)]


def build(reference: str = "/root/reference/ThirdParty/sundials") -> Path:
    """Compiles the oracle library (needs the reference sources)."""
    subprocess.run(["make", "-s", "-C", str(HERE), f"REF={reference}", "-j8"], check=True)
    return LIB


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            if Path("/root/reference/ThirdParty/sundials").exists():
                build()
            else:
                raise FileNotFoundError(f"{LIB} missing and the reference is not available")
        L = ctypes.CDLL(str(LIB))
        L.amx_oracle_simulate_batch.restype = ctypes.c_int
        L.amx_oracle_Jy.restype = ctypes.c_double
        L.amx_oracle_Jy.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.amx_oracle_dJydy.restype = ctypes.c_double
        L.amx_oracle_dJydy.argtypes = L.amx_oracle_Jy.argtypes
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class OracleModel:
    def __init__(self, network):
        self.net = network
        t = network.flat_tables()
        self._keep = t
        self.c = _Model(network.nx, network.nr, network.np, network.nk, network.ny,
                        _p(t["rptr"]), _p(t["ridx"]), _p(t["pptr"]), _p(t["pidx"]),
                        _p(t["kidx"]), _p(t["x0c"]), _p(t["x0k"]), _p(t["C"]),
                        int(network.noise))

    def xdot(self, theta, x):
        p = 10.0 ** np.asarray(theta, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros(self.net.nx)
        lib().amx_oracle_xdot(ctypes.byref(self.c), _p(p), _p(x), _p(out))
        return out

    def jac(self, theta, x):
        p = 10.0 ** np.asarray(theta, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros((self.net.nx, self.net.nx))
        lib().amx_oracle_jac(ctypes.byref(self.c), _p(p), _p(x), _p(out))
        return out

    def dxdotdp(self, theta, x):
        p = 10.0 ** np.asarray(theta, np.float64)
        x = np.ascontiguousarray(x, np.float64)
        out = np.zeros((self.net.nx, self.net.np))
        lib().amx_oracle_dxdotdp(ctypes.byref(self.c), _p(p), _p(x), _p(out))
        return out


def simulate(network, theta, kfix, ts, meas=None, sigma=None, *, rtol=1e-10, atol=1e-14,
             sensi=0, ism=CV_SIMULTANEOUS, interp=CV_POLYNOMIAL, stldet=1, maxsteps=100000,
             quad_rtol=1e-8, quad_atol=1e-12, nthreads=1, want_sx=False, kpre=None,
             t_preeq=1e6, plist=None, pscale=None, x0=None, sx0=None):
    """CVODES-oracle batch: theta [nb, np] (on the scale pscale [np] of
    ParameterScaling 0 none / 1 ln / 2 log10, default all log10), kfix [nb,
    nk], ts [nt] or [nb, nt], meas / sigma [nb, nt, ny].  sensi: 0 none, 1
    forward, 2 adjoint, with respect to the parameters plist (default all).
    kpre [nb, nk]: fixed parameters of a pre-equilibration to steady state
    (integration to t_preeq), after which the fixed-parameter-dependent
    initial states are reset from kfix.  x0 [nb, nx]: initial states
    replacing the model's; sx0 [nb, nplist, nx]: their sensitivities with
    respect to the scaled parameters (without kpre).
    Returns dict of numpy arrays x, y, sx, llh, sllh, status, counters."""
    om = OracleModel(network)
    theta = np.ascontiguousarray(np.atleast_2d(theta), np.float64)
    nb = theta.shape[0]
    kfix = np.ascontiguousarray(np.asarray(kfix, np.float64).reshape(nb, network.nk))
    ts = np.asarray(ts, np.float64)
    nt = ts.shape[-1]
    ts = np.ascontiguousarray(np.broadcast_to(ts, (nb, nt)))
    if sigma is None:
        sigma = np.ones((nb, nt, network.ny))
    sigma = np.ascontiguousarray(np.broadcast_to(sigma, (nb, nt, network.ny)), np.float64)
    meas = (np.ascontiguousarray(np.broadcast_to(meas, (nb, nt, network.ny)), np.float64)
            if meas is not None else None)
    x = np.zeros((nb, nt, network.nx))
    y = np.zeros((nb, nt, network.ny))
    pl = None if plist is None else np.ascontiguousarray(plist, np.int32)
    ps = None if pscale is None else np.ascontiguousarray(pscale, np.int32)
    npl = network.np if pl is None else len(pl)
    x0a = None if x0 is None else np.ascontiguousarray(np.asarray(x0, np.float64).reshape(nb, network.nx))
    sx0a = (None if sx0 is None
            else np.ascontiguousarray(np.asarray(sx0, np.float64).reshape(nb, npl, network.nx)))
    sx = np.zeros((nb, nt, npl, network.nx)) if (want_sx and sensi == 1) else None
    llh = np.zeros(nb)
    sllh = np.zeros((nb, npl))
    counters = np.zeros((nb, NCOUNTERS), np.int64)
    status = np.zeros(nb, np.int32)
    opts = _Opts(rtol, atol, quad_rtol, quad_atol, maxsteps, sensi, ism, interp, stldet, t_preeq,
                 npl, _p(pl), _p(ps))
    if kpre is not None:
        kpre = np.ascontiguousarray(np.asarray(kpre, np.float64).reshape(nb, network.nk))
    lib().amx_oracle_simulate_batch(
        ctypes.byref(om.c), ctypes.byref(opts), ctypes.c_int(nb), _p(theta), _p(kfix), _p(kpre),
        _p(x0a), _p(sx0a), ctypes.c_int(nt), _p(ts), _p(meas), _p(sigma), _p(x), _p(y), _p(sx), _p(llh), _p(sllh),
        _p(counters), _p(status), ctypes.c_int(nthreads))
    return dict(x=x, y=y, sx=sx, llh=llh, sllh=sllh if sensi else None, status=status,
                counters=counters)


def trace_forward(network, theta, kfix, ts, *, rtol=1e-8, atol=1e-12, stldet=1, maxsteps=100000,
                  cap=200000) -> np.ndarray:
    """Accepted steps of one instance's forward integration (states only):
    rows of t, step size used, order used, nst, nfe, nsetups, nje, netf, ncfn."""
    om = OracleModel(network)
    theta = np.ascontiguousarray(theta, np.float64)
    kfix = np.ascontiguousarray(np.asarray(kfix, np.float64).reshape(network.nk))
    ts = np.ascontiguousarray(ts, np.float64)
    rec = np.zeros((cap, 9))
    opts = _Opts(rtol, atol, 1e-8, 1e-12, maxsteps, 0, CV_SIMULTANEOUS, CV_POLYNOMIAL, stldet, 0.0,
                 0, None, None)
    L = lib()
    n = L.amx_oracle_trace_forward(ctypes.byref(om.c), ctypes.byref(opts), _p(theta), _p(kfix),
                                   ctypes.c_int(len(ts)), _p(ts), _p(rec), ctypes.c_int(cap))
    if n < 0:
        raise RuntimeError(f"oracle trace failed ({n})")
    return rec[:min(n, cap)]


STALD_LIB = HERE / "_ref" / "libamx_stald_ref.so"
_stald = None


def _stald_lib() -> ctypes.CDLL:
    global _stald
    if _stald is None:
        if not STALD_LIB.exists():
            if Path("/root/reference/ThirdParty/sundials").exists():
                build()
            else:
                raise FileNotFoundError(f"{STALD_LIB} missing and the reference is not available")
        L = ctypes.CDLL(str(STALD_LIB))
        L.amx_oracle_sldet.restype = ctypes.c_int
        L.amx_oracle_sldet.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.amx_oracle_stald_samples.restype = ctypes.c_int
        _stald = L
    return _stald


def sldet(ss, q: int) -> int:
    """CVODES's stability-limit detector (cvSLdet) on ss [5][3] (newest step
    first; orders q-1, q, q+1 of the squared scaled derivative norms)."""
    ss = np.ascontiguousarray(ss, np.float64).reshape(5, 3)
    return int(_stald_lib().amx_oracle_sldet(_p(ss), int(q)))


def stald_samples(network, theta, kfix, ts, *, rtol=1e-8, atol=1e-12, cap=50000) -> np.ndarray:
    """The detector inputs of one forward integration with stability-limit
    detection on: rows (q, ss[5][3] flattened) of every step at which CVODES
    called the detector."""
    om = OracleModel(network)
    theta = np.ascontiguousarray(theta, np.float64)
    kfix = np.ascontiguousarray(np.asarray(kfix, np.float64).reshape(network.nk))
    ts = np.ascontiguousarray(ts, np.float64)
    out = np.zeros((cap, 16))
    opts = _Opts(rtol, atol, 1e-8, 1e-12, 100000, 0, CV_SIMULTANEOUS, CV_POLYNOMIAL, 1, 0.0,
                 0, None, None)
    n = _stald_lib().amx_oracle_stald_samples(
        ctypes.byref(om.c), ctypes.byref(opts), _p(theta), _p(kfix), ctypes.c_int(len(ts)),
        _p(ts), _p(out), ctypes.c_int(cap))
    if n < 0:
        raise RuntimeError(f"oracle integration failed ({n})")
    return out[:min(n, cap)]


def Jy(noise: int, y: float, sigma: float, m: float) -> float:
    return lib().amx_oracle_Jy(noise, y, sigma, m)


def dJydy(noise: int, y: float, sigma: float, m: float) -> float:
    return lib().amx_oracle_dJydy(noise, y, sigma, m)
