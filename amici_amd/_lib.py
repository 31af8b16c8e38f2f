"""ctypes binding of the C ABI declared in include/amx.h.

The product path has no CPU fallback: loading a model module that is missing
or was built without its HIP kernels raises immediately.
"""
from __future__ import annotations

import ctypes
import math
from pathlib import Path

AMX_ABI_VERSION = 10
AMX_NSTATS = 18
AMX_NDIAG = 9
SENSI_NONE, SENSI_FORWARD, SENSI_ADJOINT = 0, 1, 2
# ParameterScaling (include/amici/defines.h)
PSCALE_NONE, PSCALE_LN, PSCALE_LOG10 = 0, 1, 2
INTERP_HERMITE, INTERP_POLYNOMIAL = 1, 2
STLDET_ON, STLDET_OFF = 1, 2

SYMBOLS = ("amx_model_info", "amx_workspace_bytes", "amx_run_batch", "amx_run_sequential",
           "amx_set_trace", "amx_set_timing", "amx_kernel_times", "amx_status_string",
           "amx_graph_create", "amx_graph_launch", "amx_graph_destroy", "amx_run_batch_static",
           "amx_static_workspace_bytes")


class SolverOpts(ctypes.Structure):
    _fields_ = [
        ("rtol", ctypes.c_double), ("atol", ctypes.c_double),
        ("rtol_fsa", ctypes.c_double), ("atol_fsa", ctypes.c_double),
        ("rtolB", ctypes.c_double), ("atolB", ctypes.c_double),
        ("quad_rtol", ctypes.c_double), ("quad_atol", ctypes.c_double),
        ("maxsteps", ctypes.c_int), ("maxstepsB", ctypes.c_int),
        ("sensi_meth", ctypes.c_int), ("max_traj", ctypes.c_int),
        ("maxtime", ctypes.c_double),
        ("newton_maxsteps", ctypes.c_int),
        ("ss_rtol", ctypes.c_double), ("ss_atol", ctypes.c_double),
        ("max_blocks", ctypes.c_int), ("interp", ctypes.c_int),
        ("stldet", ctypes.c_int), ("failfast", ctypes.c_int),
        ("ckpt_steps", ctypes.c_int),
    ]


class BatchIn(ctypes.Structure):
    _fields_ = [
        ("nb", ctypes.c_int), ("nt", ctypes.c_int), ("ts_stride", ctypes.c_int),
        ("theta", ctypes.c_void_p), ("kfix", ctypes.c_void_p), ("ts", ctypes.c_void_p),
        ("meas", ctypes.c_void_p), ("sigma", ctypes.c_void_p),
        ("kfix_preeq", ctypes.c_void_p),
        ("nplist", ctypes.c_int), ("plist", ctypes.c_void_p), ("pscale", ctypes.c_void_p),
        ("x0", ctypes.c_void_p), ("sx0", ctypes.c_void_p), ("t0", ctypes.c_void_p),
    ]


class BatchOut(ctypes.Structure):
    _fields_ = [
        ("x", ctypes.c_void_p), ("y", ctypes.c_void_p), ("sx", ctypes.c_void_p),
        ("llh", ctypes.c_void_p), ("sllh", ctypes.c_void_p),
        ("status", ctypes.c_void_p), ("stats", ctypes.c_void_p),
        ("diag", ctypes.c_void_p),
    ]


class ModelLib:
    """A loaded model module ``libamx_<name>.so``."""

    def __init__(self, path: str | Path):
        path = Path(path)
        if not path.exists():
            raise FileNotFoundError(
                f"HIP model module {path} is missing; build it with __graft_entry__.build() "
                "or amici_amd.codegen.compile_model")
        self.path = path
        self.lib = ctypes.CDLL(str(path))
        for s in SYMBOLS:
            if not hasattr(self.lib, s):
                raise RuntimeError(f"{path} does not export {s}")
        L = self.lib
        L.amx_model_info.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.amx_model_info.restype = ctypes.c_int
        L.amx_workspace_bytes.argtypes = [ctypes.POINTER(SolverOpts), ctypes.c_int, ctypes.c_int]
        L.amx_workspace_bytes.restype = ctypes.c_size_t
        for fn in (L.amx_run_batch, L.amx_run_sequential):
            fn.argtypes = [ctypes.POINTER(SolverOpts), ctypes.POINTER(BatchIn),
                           ctypes.POINTER(BatchOut), ctypes.c_void_p, ctypes.c_size_t,
                           ctypes.c_void_p]
            fn.restype = ctypes.c_int
        L.amx_set_trace.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.amx_set_trace.restype = ctypes.c_int
        L.amx_set_timing.argtypes = [ctypes.c_int]
        L.amx_set_timing.restype = ctypes.c_int
        L.amx_kernel_times.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)]
        L.amx_kernel_times.restype = ctypes.c_int
        L.amx_run_batch_static.argtypes = [ctypes.POINTER(SolverOpts), ctypes.POINTER(BatchIn),
                                           ctypes.POINTER(BatchOut), ctypes.c_void_p]
        L.amx_run_batch_static.restype = ctypes.c_int
        L.amx_static_workspace_bytes.argtypes = []
        L.amx_static_workspace_bytes.restype = ctypes.c_size_t
        L.amx_graph_create.argtypes = [ctypes.POINTER(SolverOpts), ctypes.POINTER(BatchIn),
                                       ctypes.POINTER(BatchOut), ctypes.c_void_p, ctypes.c_size_t,
                                       ctypes.POINTER(ctypes.c_void_p)]
        L.amx_graph_create.restype = ctypes.c_int
        L.amx_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.amx_graph_launch.restype = ctypes.c_int
        L.amx_graph_destroy.argtypes = [ctypes.c_void_p]
        L.amx_graph_destroy.restype = ctypes.c_int
        L.amx_status_string.argtypes = [ctypes.c_int]
        L.amx_status_string.restype = ctypes.c_char_p
        dims = (ctypes.c_int * 8)()
        if L.amx_model_info(dims) != 0:
            raise RuntimeError("amx_model_info failed")
        self.nx, self.nr, self.np, self.nk, self.ny, self.team, self.noise, abi = list(dims)
        if abi != AMX_ABI_VERSION:
            raise RuntimeError(f"{path}: ABI version {abi} != {AMX_ABI_VERSION}")

    def status_string(self, s: int) -> str:
        return self.lib.amx_status_string(int(s)).decode()

    def set_timing(self, enable: bool) -> None:
        self.lib.amx_set_timing(int(bool(enable)))

    def kernel_times(self) -> dict:
        """Summed forward / backward kernel milliseconds and launch counts since
        the previous read (HIP events on the launch stream)."""
        ms = (ctypes.c_double * 2)()
        n = (ctypes.c_int * 2)()
        dropped = self.lib.amx_kernel_times(ms, n)
        if dropped < 0:
            raise RuntimeError(f"amx_kernel_times failed ({dropped})")
        return dict(fwd_ms=ms[0], bwd_ms=ms[1], fwd_launches=n[0], bwd_launches=n[1],
                    dropped=dropped)

    def workspace_bytes(self, opts: SolverOpts, nb: int, nt: int) -> int:
        return int(self.lib.amx_workspace_bytes(ctypes.byref(opts), int(nb), int(nt)))


def make_opts(rtol=1e-8, atol=1e-16, sensi_meth=SENSI_NONE, maxsteps=10000, rtol_fsa=math.nan,
              atol_fsa=math.nan, rtolB=math.nan, atolB=math.nan, quad_rtol=1e-8,
              quad_atol=1e-12, maxstepsB=0, max_traj=0, maxtime=0.0, newton_maxsteps=40,
              ss_rtol=math.nan, ss_atol=math.nan, max_blocks=0,
              interp=INTERP_POLYNOMIAL, stldet=True, failfast=False,
              ckpt_steps=0) -> SolverOpts:
    return SolverOpts(rtol, atol, rtol_fsa, atol_fsa, rtolB, atolB, quad_rtol, quad_atol,
                      int(maxsteps), int(maxstepsB), int(sensi_meth), int(max_traj),
                      float(maxtime), int(newton_maxsteps), float(ss_rtol), float(ss_atol),
                      int(max_blocks), int(interp), STLDET_ON if stldet else STLDET_OFF,
                      1 if failfast else 0, int(ckpt_steps))
