"""Drop-in simulation API: Model / Solver / ExpData / ReturnData and
``run_simulations``.

Mirrors the reference's Python entry points for ODE models:
``amici.sim.sundials.run_simulations(model, solver, edata_list, failfast=True,
num_threads=1)`` (python/sdist/amici/sim/sundials/_swig_wrappers.py:87-133,
formerly ``runAmiciSimulations``), ``run_simulation`` (:46-84), the
``ReturnDataView`` field layout (python/sdist/amici/sim/sundials/_numpy.py:414-451)
and the Solver / Model / ExpData setters used on this path.  The whole list of
conditions is run as one batch on the GPU through the C ABI
(``include/amx.h``); ``num_threads`` is accepted and ignored.
"""
from __future__ import annotations

import enum
import math
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np
import torch

from . import _lib, rdata as _rdata
from .networks import ReactionNetwork
from .rdata import RDataReporting

AMICI_SUCCESS = 0
AMICI_NOT_RUN = -1001


class SensitivityOrder(enum.IntEnum):
    none = 0
    first = 1


class SensitivityMethod(enum.IntEnum):
    none = 0
    forward = 1
    adjoint = 2


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("amici_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


class Model:
    """A compiled mass-action model (network + HIP model module)."""

    def __init__(self, network: ReactionNetwork, library: str | None = None,
                 team: int | None = None, waves: int | None = None):
        """``team`` / ``waves``: a launch-shape variant of the model module
        (lanes per instance, register budget in waves per SIMD); the defaults
        suit most batches, and ``amici_amd.autotune.tune`` picks them by
        measurement."""
        from . import codegen

        self.network = network
        self.team, self.waves = team, waves
        path = library or codegen.compile_model(network, team, waves=waves)
        self.lib = _lib.ModelLib(path)
        if (self.lib.nx, self.lib.np, self.lib.nk, self.lib.ny) != (
                network.nx, network.np, network.nk, network.ny):
            raise RuntimeError(f"{path} was built for a different network")
        self._theta = np.asarray(network.theta_nominal, np.float64).copy()
        self._k = np.asarray(network.k_nominal, np.float64).copy()
        self._ts = np.zeros(0)
        self._ws: dict = {}
        self._sigma_res = False
        self._min_sigma = 50.0
        self._plist = list(range(network.np))
        self._pscale = [_lib.PSCALE_LOG10] * network.np
        self._x0 = None
        self._sx0 = None
        self._t0 = 0.0

    # dimensions
    @property
    def nx_rdata(self) -> int:
        return self.network.nx

    @property
    def ny(self) -> int:
        return self.network.ny

    @property
    def np(self) -> int:
        return self.network.np

    @property
    def nk(self) -> int:
        return self.network.nk

    ne = 0

    def nplist(self) -> int:
        return len(self._plist)

    # ids
    def get_state_ids(self):
        return list(self.network.species)

    def get_observable_ids(self):
        return [f"y{k}" for k in range(self.ny)]

    def get_free_parameter_ids(self):
        return list(self.network.parameter_names)

    def get_fixed_parameter_ids(self):
        return list(self.network.fixed_parameter_names)

    def get_parameter_list(self):
        return list(self._plist)

    def set_parameter_list(self, plist: Sequence[int]) -> None:
        """Parameters whose sensitivities are computed (Model::set_parameter_list,
        include/amici/model.h:954): sllh / sx columns follow this list."""
        pl = [int(i) for i in plist]
        if any(i < 0 or i >= self.np for i in pl):
            raise ValueError(f"parameter indices must be in [0, {self.np})")
        self._plist = pl

    def get_parameter_scale(self):
        return list(self._pscale)

    def set_parameter_scale(self, pscale) -> None:
        """Scale of every free parameter (Model::set_parameter_scale,
        include/amici/model.h:414-423): one ParameterScaling for all or one per
        parameter, 0 / "none", 1 / "ln", 2 / "log10"; the free parameters and the
        sensitivities are on this scale."""
        self._pscale = _pscale_list(pscale, self.np)

    def set_initial_state(self, x0) -> None:
        """Initial state replacing the model's (Model::set_initial_state);
        an empty sequence restores the model's own."""
        x0 = np.asarray(x0, np.float64).ravel()
        if x0.size == 0:
            self._x0 = None
            return
        if x0.shape != (self.nx_rdata,):
            raise ValueError(f"expected {self.nx_rdata} initial states")
        self._x0 = x0.copy()

    def set_t0(self, t0: float) -> None:
        """Start time of the simulation (Model::set_t0; ExpData.t_start
        overrides it per condition, src/edata.cpp:551)."""
        self._t0 = float(t0)

    def t0(self) -> float:
        return self._t0

    def set_initial_state_sensitivities(self, sx0) -> None:
        """Initial state sensitivities with respect to the scaled parameters
        of the parameter list (Model::set_initial_state_sensitivities,
        src/model.cpp:986-1023; nplist x nx, AMICI's row-major order); an
        empty sequence restores the model's own (zero)."""
        sx0 = np.asarray(sx0, np.float64).ravel()
        if sx0.size == 0:
            self._sx0 = None
            return
        if sx0.size != self.nplist() * self.nx_rdata:
            raise ValueError(f"expected {self.nplist()} x {self.nx_rdata} initial state sensitivities")
        self._sx0 = sx0.reshape(self.nplist(), self.nx_rdata).copy()

    def get_initial_state_sensitivities(self) -> np.ndarray:
        if self._sx0 is not None:
            return self._sx0.copy()
        return np.zeros((self.nplist(), self.nx_rdata))

    def get_initial_state(self, kfix=None) -> np.ndarray:
        """The initial state used: the one set, else the model's (constants
        and fixed parameters)."""
        if self._x0 is not None:
            return self._x0.copy()
        net = self.network
        k = self._k if kfix is None else np.asarray(kfix, np.float64)
        return np.array([k[j] if j >= 0 else c for j, c in zip(net.x0_fixed, net.x0_const)],
                        np.float64)

    # values (free parameters on their scale, as AMICI stores them)
    def set_free_parameters(self, theta: Sequence[float]) -> None:
        theta = np.asarray(theta, np.float64)
        if theta.shape != (self.np,):
            raise ValueError(f"expected {self.np} free parameters")
        self._theta = theta.copy()

    def get_free_parameters(self) -> np.ndarray:
        return self._theta.copy()

    def set_fixed_parameters(self, k: Sequence[float]) -> None:
        k = np.asarray(k, np.float64)
        if k.shape != (self.nk,):
            raise ValueError(f"expected {self.nk} fixed parameters")
        self._k = k.copy()

    def get_fixed_parameters(self) -> np.ndarray:
        return self._k.copy()

    def set_timepoints(self, ts: Sequence[float]) -> None:
        self._ts = np.asarray(ts, np.float64).copy()

    def get_timepoints(self) -> np.ndarray:
        return self._ts.copy()

    def nt(self) -> int:
        return len(self._ts)

    # error residuals sqrt(2 log sigma + min_sigma) appended to res
    # (Model::set_add_sigma_residuals, include/amici/model.h:885-929)
    def set_add_sigma_residuals(self, v: bool) -> None:
        self._sigma_res = bool(v)

    def get_add_sigma_residuals(self) -> bool:
        return self._sigma_res

    def set_minimum_sigma_residuals(self, v: float) -> None:
        self._min_sigma = float(v)

    def get_minimum_sigma_residuals(self) -> float:
        return self._min_sigma

    def create_solver(self) -> "Solver":
        return Solver()

    def simulate(self, *, solver: "Solver | None" = None, edata=None, failfast: bool = True,
                 num_threads: int = 1, sensi_method=None, sensi_order=None):
        """``Model.simulate`` (python/sdist/amici/sim/sundials/_swig_wrappers.py:296-331):
        one ExpData (or none) -> one ReturnData, a sequence -> a list.  The
        sensitivity method / order ("forward", "adjoint", "first", or the
        enums) apply to this call only; the solver's own settings are restored."""
        solver = solver if solver is not None else self.create_solver()
        old = (solver.sensi_meth, solver.sensi)
        try:
            if sensi_method is not None:
                solver.set_sensitivity_method(SensitivityMethod[sensi_method]
                                              if isinstance(sensi_method, str) else sensi_method)
            if sensi_order is not None:
                solver.set_sensitivity_order(SensitivityOrder[sensi_order]
                                             if isinstance(sensi_order, str) else sensi_order)
            if isinstance(edata, Sequence):
                return run_simulations(self, solver, edata, failfast=failfast, num_threads=num_threads)
            return run_simulation(self, solver, edata)
        finally:
            solver.sensi_meth, solver.sensi = old

    def workspace(self, nbytes: int, device, stream) -> torch.Tensor:
        """Device workspace of the calls on ``stream`` (work-queue counters,
        adjoint trajectory): one per stream, so calls in flight on different
        streams never share one, while calls on one stream run in order.  The
        buffer is allocated on that stream, so the caching allocator orders
        its reuse after them."""
        key = (str(device), stream.cuda_stream)
        ws = self._ws.get(key)
        if ws is None or ws.numel() < nbytes:
            with torch.cuda.stream(stream):
                ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
            self._ws[key] = ws
        return ws


@dataclass
class Solver:
    """CVODES-equivalent solver options with AMICI's defaults
    (include/amici/solver.h:1869-2147)."""

    rtol: float = 1e-8
    atol: float = 1e-16
    rtol_fsa: float = math.nan
    atol_fsa: float = math.nan
    rtolB: float = math.nan
    atolB: float = math.nan
    quad_rtol: float = 1e-8
    quad_atol: float = 1e-12
    maxsteps: int = 10000
    maxstepsB: int = 0
    sensi: SensitivityOrder = SensitivityOrder.none
    sensi_meth: SensitivityMethod = SensitivityMethod.forward
    max_traj: int = 0
    maxtime: float = 0.0
    newton_maxsteps: int = 40
    ss_rtol: float = math.nan
    ss_atol: float = math.nan
    rdata_reporting: RDataReporting = RDataReporting.full
    max_blocks: int = 0  # launch shape: 0 = as many blocks as fit the device at once
    interp: int = 2      # InterpolationType: 1 hermite, 2 polynomial (AMICI's default)
    stldet: bool = True  # BDF stability-limit detection (AMICI's default: on)
    ckpt_steps: int = 0  # adjoint: checkpoint every K forward steps (0: store every step)

    def set_relative_tolerance(self, v): self.rtol = float(v)
    def get_relative_tolerance(self): return self.rtol
    def set_absolute_tolerance(self, v): self.atol = float(v)
    def get_absolute_tolerance(self): return self.atol
    def set_relative_tolerance_fsa(self, v): self.rtol_fsa = float(v)
    def set_absolute_tolerance_fsa(self, v): self.atol_fsa = float(v)
    def set_relative_tolerance_b(self, v): self.rtolB = float(v)
    def set_absolute_tolerance_b(self, v): self.atolB = float(v)
    def set_relative_tolerance_quadratures(self, v): self.quad_rtol = float(v)
    def set_absolute_tolerance_quadratures(self, v): self.quad_atol = float(v)
    def set_max_steps(self, v): self.maxsteps = int(v)
    def get_max_steps(self): return self.maxsteps
    def set_max_steps_backward_problem(self, v): self.maxstepsB = int(v)
    def set_max_time(self, v): self.maxtime = float(v)
    def get_max_time(self): return self.maxtime
    def set_newton_max_steps(self, v): self.newton_maxsteps = int(v)
    def get_newton_max_steps(self): return self.newton_maxsteps
    def set_relative_tolerance_steady_state(self, v): self.ss_rtol = float(v)
    def set_absolute_tolerance_steady_state(self, v): self.ss_atol = float(v)
    def set_return_data_reporting_mode(self, v): self.rdata_reporting = RDataReporting(int(v))
    def get_return_data_reporting_mode(self): return self.rdata_reporting
    def set_sensitivity_order(self, v): self.sensi = SensitivityOrder(int(v))
    def get_sensitivity_order(self): return self.sensi
    def set_sensitivity_method(self, v): self.sensi_meth = SensitivityMethod(int(v))
    def get_sensitivity_method(self): return self.sensi_meth

    def set_interpolation_type(self, v):
        """Interpolation of the forward solution in the adjoint sweep
        (Solver::setInterpolationType; InterpolationType hermite = 1,
        polynomial = 2, the default)."""
        name = getattr(v, "name", str(v)).lower()
        if name in ("hermite", "1"):
            self.interp = 1
        elif name in ("polynomial", "2"):
            self.interp = 2
        else:
            raise ValueError(f"unknown interpolation type {v!r}")

    def get_interpolation_type(self):
        return self.interp

    def set_stability_limit_flag(self, v):
        """BDF stability-limit detection in the forward and backward problems
        (Solver::set_stability_limit_flag, include/amici/solver.h:664-673; on
        by default)."""
        self.stldet = bool(v)

    def get_stability_limit_flag(self):
        return self.stldet

    def set_checkpoint_steps(self, v):
        """Adjoint checkpoints every ``v`` forward steps (CVodeAdjInit(Nd),
        cvodea.c:99): the forward pass keeps a checkpoint of the integrator
        state every ``v`` steps instead of every step, and the backward sweep
        recomputes each interval from its checkpoint.  0 (the default) stores
        every step, which is what AMICI does (CVodeAdjInit(maxsteps + 1),
        src/solver_cvodes.cpp:726-739).  The integrator takes a fresh Newton
        matrix at each checkpoint, so results differ from the default within
        the integration tolerances."""
        if int(v) < 0:
            raise ValueError("checkpoint steps must be >= 0")
        self.ckpt_steps = int(v)

    def get_checkpoint_steps(self):
        return self.ckpt_steps

    def set_max_blocks(self, v):
        """Caps the blocks of the persistent kernels (no counterpart in the
        reference): with fewer blocks than instances each block takes further
        instances from the device work queue.  0 restores the default."""
        if int(v) < 0:
            raise ValueError("max_blocks must be >= 0")
        self.max_blocks = int(v)

    def set_linear_solver(self, v):
        if str(v).lower() not in ("dense", "linearsolver.dense", "1"):
            raise ValueError("this backend uses the dense direct linear solver")

    def effective_method(self) -> int:
        if self.sensi == SensitivityOrder.none:
            return _lib.SENSI_NONE
        return int(self.sensi_meth)

    def c_opts(self) -> _lib.SolverOpts:
        return _lib.make_opts(rtol=self.rtol, atol=self.atol, sensi_meth=self.effective_method(),
                              maxsteps=self.maxsteps, rtol_fsa=self.rtol_fsa,
                              atol_fsa=self.atol_fsa, rtolB=self.rtolB, atolB=self.atolB,
                              quad_rtol=self.quad_rtol, quad_atol=self.quad_atol,
                              maxstepsB=self.maxstepsB, max_traj=self.max_traj,
                              maxtime=self.maxtime, newton_maxsteps=self.newton_maxsteps,
                              ss_rtol=self.ss_rtol, ss_atol=self.ss_atol,
                              max_blocks=self.max_blocks, interp=self.interp,
                              stldet=self.stldet, ckpt_steps=self.ckpt_steps)


class ExpData:
    """Per-condition inputs (include/amici/edata.h): time grid, measurements
    [nt, ny] (NaN = missing), noise scales [nt, ny], fixed parameters, fixed
    parameters of a pre-equilibration (empty = none), and the optional
    overrides of the model's settings that ConditionContext::apply_condition
    applies (src/edata.cpp:438-490): free parameters (on their scale),
    parameter list ``plist``, parameter scales ``pscale``, initial states
    ``x0`` and their sensitivities ``sx0`` (nplist x nx, with respect to the
    scaled parameters; empty = the model's), and the start time ``t_start``
    (None = the model's t0)."""

    def __init__(self, source: "Model | ReturnData | None" = None, sigma_y=None,
                 sigma_z: float = 0.0, seed: int = -1):
        """``ExpData(model)``: empty data on the model's time grid.
        ``ExpData(rdata, sigma_y, sigma_z, seed)``: synthetic data from a
        simulation, y + N(0, sigma_y) with noise scales sigma_y (scalar,
        [ny] or [nt, ny]; src/edata.cpp:74-134; seed < 0 draws a random
        seed).  There are no event observables, so sigma_z is unused."""
        del sigma_z
        if isinstance(source, ReturnData):
            if source.y is None:
                raise ValueError("the ReturnData holds no observables")
            self.ny = source.y.shape[1]
            self.timepoints = np.asarray(source.ts, np.float64).copy()
            self.fixed_parameters = np.zeros(0)
            self.free_parameters = np.zeros(0)
            self.fixed_parameters_pre_equilibration = np.zeros(0)
            self.fixed_parameters_presimulation = np.zeros(0)
            self.t_presim = 0.0
            self.plist, self.pscale, self.x0, self.sx0 = [], [], np.zeros(0), np.zeros(0)
            self.t_start = None
            nt = len(self.timepoints)
            sig = np.broadcast_to(np.asarray(1.0 if sigma_y is None else sigma_y, np.float64),
                                  (nt, self.ny)).copy()
            if not np.all(sig > 0):
                raise ValueError("sigma_y must be positive")
            rng = np.random.default_rng(None if seed < 0 else seed)
            self._meas = source.y + sig * rng.standard_normal((nt, self.ny))
            self._sigma = sig
            self.id = ""
            return
        model = source
        self.ny = model.ny if model else 0
        self.timepoints = model.get_timepoints() if model else np.zeros(0)
        self.fixed_parameters = model.get_fixed_parameters() if model else np.zeros(0)
        self.free_parameters = np.zeros(0)
        self.fixed_parameters_pre_equilibration = np.zeros(0)
        # pre-simulation over t_presim before the start time with these fixed
        # parameters (empty: the condition's)
        self.fixed_parameters_presimulation = np.zeros(0)
        self.t_presim = 0.0
        self.plist, self.pscale, self.x0, self.sx0 = [], [], np.zeros(0), np.zeros(0)
        self.t_start = None  # start time (None: the model's t0)
        self._meas = np.full((len(self.timepoints), self.ny), np.nan)
        self._sigma = np.ones((len(self.timepoints), self.ny))
        self.id = ""

    def set_timepoints(self, ts):
        self.timepoints = np.asarray(ts, np.float64).copy()
        self._meas = np.full((len(self.timepoints), self.ny), np.nan)
        self._sigma = np.ones((len(self.timepoints), self.ny))

    def get_timepoints(self):
        return self.timepoints.copy()

    def set_measurements(self, m):
        self._meas = np.asarray(m, np.float64).reshape(len(self.timepoints), self.ny).copy()

    def get_measurements(self):
        return self._meas.copy()

    def set_noise_scales(self, s):
        s = np.asarray(s, np.float64)
        if s.ndim == 0:
            s = np.full((len(self.timepoints), self.ny), float(s))
        self._sigma = s.reshape(len(self.timepoints), self.ny).copy()

    def get_noise_scales(self):
        return self._sigma.copy()


@dataclass
class ReturnData:
    """Results of one condition (fields and shapes of ReturnDataView,
    python/sdist/amici/sim/sundials/_numpy.py:412-450); fields the reporting
    mode leaves out are None."""

    ts: np.ndarray
    x: np.ndarray | None = None
    y: np.ndarray | None = None
    sigmay: np.ndarray | None = None
    sx: np.ndarray | None = None
    sy: np.ndarray | None = None
    ssigmay: np.ndarray | None = None
    llh: float = math.nan
    sllh: np.ndarray | None = None
    chi2: float = math.nan
    res: np.ndarray | None = None
    sres: np.ndarray | None = None
    FIM: np.ndarray | None = None
    sigma_res: bool = False
    status: int = AMICI_NOT_RUN
    # per-output-time solver diagnostics (include/amici/rdata.h:261-301):
    # cumulative forward counters at each output time, and the backward
    # problem's cumulative counters in the order the backward sweep stores
    # them (entry k at the k-th data point from the last one)
    numsteps: np.ndarray | None = None
    num_rhs_evals: np.ndarray | None = None
    num_err_test_fails: np.ndarray | None = None
    num_non_lin_solv_conv_fails: np.ndarray | None = None
    order: np.ndarray | None = None
    numsteps_b: np.ndarray | None = None
    num_rhs_evals_b: np.ndarray | None = None
    num_err_test_fails_b: np.ndarray | None = None
    num_non_lin_solv_conv_fails_b: np.ndarray | None = None
    stats: np.ndarray = field(default_factory=lambda: np.zeros(_lib.AMX_NSTATS, np.int64))
    messages: list = field(default_factory=list)

    def __getitem__(self, key):
        return getattr(self, key)


DIAG_FIELDS = ("numsteps", "num_rhs_evals", "num_err_test_fails", "num_non_lin_solv_conv_fails",
               "order", "numsteps_b", "num_rhs_evals_b", "num_err_test_fails_b",
               "num_non_lin_solv_conv_fails_b")


def _pscale_list(pscale, n: int) -> list:
    names = {"none": _lib.PSCALE_NONE, "ln": _lib.PSCALE_LN, "log10": _lib.PSCALE_LOG10}

    def one(v):
        key = getattr(v, "name", v)
        if isinstance(key, str):
            if key.lower() not in names:
                raise ValueError(f"unknown parameter scale {v!r}")
            return names[key.lower()]
        if int(v) not in (0, 1, 2):
            raise ValueError(f"unknown parameter scale {v!r}")
        return int(v)

    if isinstance(pscale, (str, int)) or hasattr(pscale, "name"):
        return [one(pscale)] * n
    out = [one(v) for v in pscale]
    if len(out) != n:
        raise ValueError(f"expected {n} parameter scales")
    return out


def _as_dev(a, dev, dtype=torch.float64) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype).contiguous()
    a = np.ascontiguousarray(a)
    if not a.flags.writeable:  # broadcast views: torch wants writable memory
        a = a.copy()
    return torch.as_tensor(a, dtype=dtype, device=dev)


def _ptr(t):
    return None if t is None else t.data_ptr()


def _prepare(model: Model, solver: Solver, theta, kfix, ts, meas, sigma, *, want_x, want_y,
             want_sx, want_stats, kfix_preeq, plist=None, pscale=None, x0=None,
             sx0=None, t0=None, failfast=False, copy_inputs=False) -> dict:
    """Device inputs, output tensors and the C structs of one batch call.
    With ``copy_inputs`` every input is a fresh device copy (for a graph that
    keeps pointing at them).  plist / pscale default to the model's settings
    (Model.set_parameter_list / set_parameter_scale), x0 [nb, nx] to the
    model's initial state (Model.set_initial_state, else its own)."""
    dev = _device()
    plist = model.get_parameter_list() if plist is None else [int(i) for i in plist]
    pscale = model.get_parameter_scale() if pscale is None else _pscale_list(pscale, model.np)
    if any(i < 0 or i >= model.np for i in plist):
        raise ValueError(f"parameter indices must be in [0, {model.np})")
    npl = len(plist)
    if x0 is None and model._x0 is not None:
        x0 = np.broadcast_to(model._x0, (np.shape(theta)[0], model.nx_rdata))
    if sx0 is None and model._sx0 is not None and model._sx0.shape[0] == npl:
        sx0 = np.broadcast_to(model._sx0, (np.shape(theta)[0], npl, model.nx_rdata))
    if t0 is None:
        t0 = model._t0
    t0 = np.broadcast_to(np.asarray(t0.cpu() if isinstance(t0, torch.Tensor) else t0, np.float64),
                         (np.shape(theta)[0],))
    if not np.all(np.isfinite(t0)):
        raise ValueError("start times must be finite")
    t0 = None if not np.any(t0) else t0  # all 0: the kernels' default

    def dev_t(a):
        t = _as_dev(a, dev)
        return t.clone() if copy_inputs else t

    theta_t = dev_t(theta)
    nb = theta_t.shape[0]
    kfix_t = dev_t(kfix if kfix is not None else np.zeros((nb, model.nk)))
    ts_t = dev_t(ts)
    nt = ts_t.shape[-1]
    ts_stride = nt if ts_t.dim() == 2 else 0
    if sigma is None:
        sigma = np.ones((nb, nt, model.ny))
    sigma_t = dev_t(sigma)
    meas_t = dev_t(meas) if meas is not None else None
    kpre_t = dev_t(kfix_preeq) if kfix_preeq is not None else None
    x0_t = dev_t(x0) if x0 is not None else None
    if x0_t is not None and x0_t.shape != (nb, model.nx_rdata):
        raise ValueError("x0 must be [nb, nx]")
    sx0_t = dev_t(sx0) if sx0 is not None else None
    t0_t = dev_t(t0) if t0 is not None else None
    if sx0_t is not None and sx0_t.shape != (nb, npl, model.nx_rdata):
        raise ValueError("sx0 must be [nb, nplist, nx]")
    # the identity list and all-log10 scales are the kernels' defaults (NULL)
    plist_t = (None if plist == list(range(model.np))
               else torch.tensor(plist, dtype=torch.int32, device=dev))
    pscale_t = (None if all(v == _lib.PSCALE_LOG10 for v in pscale)
                else torch.tensor(pscale, dtype=torch.int32, device=dev))
    if kpre_t is not None and kpre_t.shape != (nb, model.nk):
        raise ValueError("kfix_preeq must be [nb, nk]")
    if theta_t.shape != (nb, model.np) or kfix_t.shape != (nb, model.nk):
        raise ValueError("theta / kfix shapes")
    if sigma_t.shape != (nb, nt, model.ny) or (meas_t is not None and meas_t.shape != sigma_t.shape):
        raise ValueError("meas / sigma shapes must be [nb, nt, ny]")
    if ts_stride and ts_t.shape[0] != nb:
        raise ValueError("ts must be [nt] or [nb, nt]")
    meth = solver.effective_method()
    opts = solver.c_opts()
    opts.failfast = 1 if failfast else 0
    f64 = dict(dtype=torch.float64, device=dev)
    out = {
        "x": torch.empty((nb, nt, model.nx_rdata), **f64) if want_x else None,
        "y": torch.empty((nb, nt, model.ny), **f64) if want_y else None,
        "sx": (torch.empty((nb, nt, npl, model.nx_rdata), **f64)
               if (want_sx and meth == _lib.SENSI_FORWARD) else None),
        "llh": torch.empty(nb, **f64),
        "sllh": torch.empty((nb, npl), **f64) if meth != _lib.SENSI_NONE else None,
        "status": torch.full((nb,), AMICI_NOT_RUN, dtype=torch.int32, device=dev),
        "stats": (torch.zeros((nb, _lib.AMX_NSTATS), dtype=torch.int32, device=dev)
                  if want_stats else None),
        "diag": (torch.zeros((nb, nt, _lib.AMX_NDIAG), dtype=torch.int32, device=dev)
                 if want_stats else None),
    }
    bin_ = _lib.BatchIn(nb, nt, ts_stride, _ptr(theta_t), _ptr(kfix_t) if model.nk else None,
                        _ptr(ts_t), _ptr(meas_t), _ptr(sigma_t), _ptr(kpre_t),
                        npl, _ptr(plist_t), _ptr(pscale_t), _ptr(x0_t), _ptr(sx0_t),
                        _ptr(t0_t))
    bout = _lib.BatchOut(*[_ptr(out[k]) for k in ("x", "y", "sx", "llh", "sllh", "status", "stats",
                                                  "diag")])
    inputs = dict(theta=theta_t, kfix=kfix_t, ts=ts_t, meas=meas_t, sigma=sigma_t,
                  kfix_preeq=kpre_t, plist=plist_t, pscale=pscale_t, x0=x0_t, sx0=sx0_t,
                  t0=t0_t)
    return dict(dev=dev, nb=nb, nt=nt, opts=opts, bin=bin_, bout=bout, out=out, inputs=inputs)


def simulate_batch(model: Model, solver: Solver, theta, kfix, ts, meas=None, sigma=None, *,
                   want_x: bool = True, want_y: bool = True, want_sx: bool = False,
                   want_stats: bool = True, sequential: bool = False, stream=None,
                   kfix_preeq=None, plist=None, pscale=None, x0=None, sx0=None, t0=None,
                   failfast: bool = False) -> dict:
    """Simulates ``nb`` instances in one batch; arrays may be numpy or torch.

    theta [nb, np] (on the parameter scales), kfix [nb, nk], ts [nt] or
    [nb, nt], meas / sigma [nb, nt, ny].  Returns a dict of device tensors: x,
    y, sx, llh, sllh, status, stats, diag.  ``sequential`` runs the plain
    one-instance-per-launch path.  ``kfix_preeq`` [nb, nk] pre-equilibrates
    every instance to steady state with these fixed parameters before the
    simulation.  ``plist`` (parameter indices; sllh [nb, len(plist)], sx [nb,
    nt, len(plist), nx]), ``pscale`` (one ParameterScaling per parameter) and
    ``x0`` [nb, nx] and ``sx0`` [nb, nplist, nx] (with respect to the scaled
    parameters) and ``t0`` (start times, scalar or [nb]) default to the
    model's settings.  With ``failfast`` the
    kernels skip instances with a larger index than a failed one that have
    not started yet (status AMICI_NOT_RUN).
    """
    import ctypes

    p = _prepare(model, solver, theta, kfix, ts, meas, sigma, want_x=want_x, want_y=want_y,
                 want_sx=want_sx, want_stats=want_stats, kfix_preeq=kfix_preeq, plist=plist,
                 pscale=pscale, x0=x0, sx0=sx0, t0=t0, failfast=failfast)
    lib, dev, opts = model.lib, p["dev"], p["opts"]
    nbytes = lib.workspace_bytes(opts, 1 if sequential else p["nb"], p["nt"])
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    ws = model.workspace(nbytes, dev, st)
    fn = lib.lib.amx_run_sequential if sequential else lib.lib.amx_run_batch
    rc = fn(ctypes.byref(opts), ctypes.byref(p["bin"]), ctypes.byref(p["bout"]), ws.data_ptr(),
            ws.numel(), ctypes.c_void_p(st.cuda_stream))
    if rc != 0:
        raise RuntimeError(f"amx_run_batch returned {rc}")
    out = p["out"]
    i = p["inputs"]
    out["_keepalive"] = tuple(i.values())
    return out


class BatchGraph:
    """One batch captured as a HIP graph (``amx_graph_create``) and replayed
    with :meth:`launch`: the launch sequence of :func:`simulate_batch` bound to
    device buffers the graph owns.  Inputs are updated in place with
    :meth:`set_inputs` (same shapes); every replay overwrites ``out``, the
    dict of output tensors.  Replaces the per-call host work of
    run_simulations (src/amici.cpp:235-262) for repeated evaluations of one
    batch shape, such as an optimiser's objective."""

    def __init__(self, model: Model, solver: Solver, theta, kfix, ts, meas=None, sigma=None, *,
                 want_x: bool = True, want_y: bool = True, want_sx: bool = False,
                 want_stats: bool = True, kfix_preeq=None, plist=None, pscale=None, x0=None,
                 sx0=None, t0=None):
        import ctypes

        p = _prepare(model, solver, theta, kfix, ts, meas, sigma, want_x=want_x, want_y=want_y,
                     want_sx=want_sx, want_stats=want_stats, kfix_preeq=kfix_preeq,
                     plist=plist, pscale=pscale, x0=x0, sx0=sx0, t0=t0, copy_inputs=True)
        self.model = model
        self.inputs = p["inputs"]
        self.out = p["out"]
        self._opts, self._bin, self._bout = p["opts"], p["bin"], p["bout"]
        nbytes = model.lib.workspace_bytes(p["opts"], p["nb"], p["nt"])
        self._ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=p["dev"])
        torch.cuda.synchronize(p["dev"])  # inputs staged before the capture reads their addresses
        h = ctypes.c_void_p()
        rc = model.lib.lib.amx_graph_create(ctypes.byref(self._opts), ctypes.byref(self._bin),
                                            ctypes.byref(self._bout), self._ws.data_ptr(),
                                            self._ws.numel(), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"amx_graph_create returned {rc}")
        self._h = h

    def set_inputs(self, **kw) -> None:
        """Copies new values into the graph's input buffers (theta, kfix, ts,
        meas, sigma, kfix_preeq), on the current stream."""
        for k, v in kw.items():
            dst = self.inputs.get(k)
            if dst is None:
                raise ValueError(f"the graph has no input {k!r}")
            src = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v))
            dst.copy_(src.reshape(dst.shape))

    def launch(self, stream=None) -> dict:
        """Replays the batch on ``stream`` (default: the current stream) and
        returns ``out`` (valid once the stream has run it)."""
        import ctypes

        if self._h is None:
            raise RuntimeError("the graph was closed")
        st = stream if stream is not None else torch.cuda.current_stream()
        rc = self.model.lib.lib.amx_graph_launch(self._h, ctypes.c_void_p(st.cuda_stream))
        if rc != 0:
            raise RuntimeError(f"amx_graph_launch returned {rc}")
        return self.out

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            self.model.lib.lib.amx_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def simulate_concurrent(groups: Sequence[dict]) -> list:
    """Several batches, typically of different models, at the same time: each
    group (the keyword arguments of :func:`simulate_batch`: model, solver,
    theta, kfix, ts and optionally meas, sigma, kfix_preeq, want_*) runs on
    its own HIP stream (default priority), so the kernels of different models
    share the GPU; inputs are staged and outputs allocated on the group's
    stream.  Waits for all groups and returns their result dicts in order.
    Results equal those of running the groups one after another."""
    dev = _device()
    streams = [torch.cuda.Stream(device=dev) for _ in groups]
    cur = torch.cuda.current_stream(dev)
    outs = []
    for g, st in zip(groups, streams):
        st.wait_stream(cur)
        with torch.cuda.stream(st):
            kw = dict(g)
            outs.append(simulate_batch(kw.pop("model"), kw.pop("solver"), kw.pop("theta"),
                                       kw.pop("kfix"), kw.pop("ts"), stream=st, **kw))
    for st in streams:
        cur.wait_stream(st)
    torch.cuda.synchronize(dev)
    return outs


def condition_start(model: Model, edata_list) -> tuple:
    """Fixed parameters [nb, nk] and start times [nb] of the conditions as the
    kernels run them.  The start time is ExpData t_start, else the model's t0
    (ConditionContext::apply_condition, src/edata.cpp:551).  Pre-simulation
    (ForwardProblem::handle_presimulation, src/forwardproblem.cpp:284-321)
    starts t_presim earlier with the pre-simulation's fixed parameters and
    continues through t0 without reinitialisation (the reference refuses
    reinitialised states there), the adjoint sweep running back through it
    (backwardproblem.cpp:53-64).  The fixed parameters of these models only
    set the initial amounts of input species (and the species reset after a
    pre-equilibration), so a pre-simulation is the condition run from
    t0 - t_presim with the pre-simulation's fixed parameters."""
    kfix, t0 = [], []
    for e in edata_list:
        pre = float(getattr(e, "t_presim", 0.0)) > 0
        k = e.fixed_parameters
        if pre and len(e.fixed_parameters_presimulation):
            k = e.fixed_parameters_presimulation
        kfix.append(np.asarray(k, np.float64) if len(k) else model.get_fixed_parameters())
        start = model.t0() if e.t_start is None else float(e.t_start)
        t0.append(start - (float(e.t_presim) if pre else 0.0))
    return np.stack(kfix).reshape(len(edata_list), model.nk), np.array(t0)


def run_simulations(model: Model, solver: Solver, edata_list: Sequence[ExpData],
                    failfast: bool = True, num_threads: int = 1) -> list:
    """All conditions of ``edata_list`` as one GPU batch; returns one
    :class:`ReturnData` per condition.  With ``failfast`` every condition after
    the first failed one is reported as not run (AMICI_NOT_RUN), as the serial
    reference loop does (src/amici.cpp:246-260)."""
    del num_threads
    if len(edata_list) == 0:
        return []
    nb = len(edata_list)

    def key(e):
        # one batch per time-grid length, parameter list and scales, and
        # whether the initial state is given (ConditionContext::apply_condition,
        # src/edata.cpp:438-490)
        pl = tuple(int(i) for i in e.plist) if len(e.plist) else tuple(model.get_parameter_list())
        ps = (tuple(_pscale_list(e.pscale, model.np)) if len(e.pscale)
              else tuple(model.get_parameter_scale()))
        return (len(e.timepoints), pl, ps, len(e.x0) > 0, len(e.sx0) > 0)

    keys = [key(e) for e in edata_list]
    if len(set(keys)) != 1:
        res = [None] * nb
        for k in sorted(set(keys)):
            idx = [i for i, kk in enumerate(keys) if kk == k]
            sub = run_simulations(model, solver, [edata_list[i] for i in idx], failfast=False)
            for i, r in zip(idx, sub):
                res[i] = r
        return _apply_failfast(res, failfast)
    _, plist, pscale, has_x0, has_sx0 = keys[0]
    theta = np.stack([e.free_parameters if len(e.free_parameters) else model.get_free_parameters()
                      for e in edata_list])
    kfix, t0 = condition_start(model, edata_list)
    ts = np.stack([e.timepoints for e in edata_list])
    meas = np.stack([e.get_measurements() for e in edata_list])
    sigma = np.stack([e.get_noise_scales() for e in edata_list])
    if np.all(ts == ts[0:1]):
        ts = ts[0]
    npre = {len(e.fixed_parameters_pre_equilibration) for e in edata_list}
    if npre != {0} and npre != {model.nk}:
        raise ValueError("pre-equilibration must be set for all conditions or for none "
                         "(run the two groups separately)")
    kpre = (np.stack([np.asarray(e.fixed_parameters_pre_equilibration, np.float64)
                      for e in edata_list]).reshape(nb, model.nk) if npre == {model.nk} and model.nk
            else None)
    meth = solver.effective_method()
    fsa = meth == _lib.SENSI_FORWARD
    fields = _rdata.reported_fields(solver.rdata_reporting, forward=fsa,
                                    sensi=meth != _lib.SENSI_NONE)
    x0 = (np.stack([np.asarray(e.x0, np.float64).reshape(model.nx_rdata) for e in edata_list])
          if has_x0 else None)
    sx0 = (np.stack([np.asarray(e.sx0, np.float64).reshape(len(plist), model.nx_rdata)
                     for e in edata_list]) if has_sx0 else None)
    out = simulate_batch(model, solver, theta, kfix, ts, meas, sigma, want_x="x" in fields,
                         want_sx=fsa, kfix_preeq=kpre, plist=list(plist), pscale=list(pscale),
                         x0=x0, sx0=sx0, t0=t0, failfast=failfast)
    dev = out["llh"].device
    meas_t, sigma_t = out["_keepalive"][3], out["_keepalive"][4]
    sy_t = None
    if out["sx"] is not None:
        C = torch.as_tensor(model.network.observable_matrix, dtype=torch.float64, device=dev)
        sy_t = out["sx"] @ C.T                                     # [nb, nt, nplist, ny]
    derived = _rdata.derived_outputs(out["y"], sy_t, meas_t, sigma_t, out["status"],
                                     model.network.noise, sigma_res=model.get_add_sigma_residuals(),
                                     sigma_offset=model.get_minimum_sigma_residuals(), fields=fields)
    torch.cuda.synchronize()
    host = {k: (v.cpu().numpy() if isinstance(v, torch.Tensor) else None)
            for k, v in out.items() if k != "_keepalive"}
    host.update({k: v.cpu().numpy() for k, v in derived.items()})
    if sy_t is not None:
        host["sy"] = sy_t.cpu().numpy()
    res = []
    for b, e in enumerate(edata_list):
        st = host["stats"][b]
        dg = host["diag"][b].astype(np.int64)
        rd = ReturnData(ts=np.asarray(e.timepoints), status=int(host["status"][b]),
                        stats=st.astype(np.int64), sigma_res=model.get_add_sigma_residuals(),
                        **{k: dg[:, i].copy() for i, k in enumerate(DIAG_FIELDS)})
        if "x" in fields:
            rd.x = host["x"][b]
        if "y" in fields:
            rd.y = host["y"][b]
            rd.sigmay = sigma[b]
        if "llh" in fields:
            rd.llh = float(host["llh"][b])
        if "sllh" in fields and host["sllh"] is not None:
            rd.sllh = host["sllh"][b]
        if "sx" in fields:
            rd.sx = host["sx"][b]
        if "sy" in fields:
            rd.sy = host["sy"][b]
            rd.ssigmay = host["ssigmay"][b]
        if "res" in fields:
            rd.res = host["res"][b]
            rd.chi2 = float(host["chi2"][b])
        if "sres" in fields:
            rd.sres = host["sres"][b]
        if "FIM" in fields:
            rd.FIM = host["FIM"][b]
        if rd.status != AMICI_SUCCESS:
            rd.messages.append(model.lib.status_string(rd.status))
        res.append(rd)
    return _apply_failfast(res, failfast)


def _apply_failfast(res: list, failfast: bool) -> list:
    if not failfast:
        return res
    failed = False
    for i, r in enumerate(res):
        if failed:
            res[i] = ReturnData(ts=r.ts, status=AMICI_NOT_RUN)
        elif r.status != AMICI_SUCCESS:
            failed = True
    return res


def run_simulation(model: Model, solver: Solver, edata: ExpData | None = None) -> ReturnData:
    """Single condition (``amici.sim.sundials.run_simulation``)."""
    if edata is None:
        edata = ExpData(model)
    return run_simulations(model, solver, [edata], failfast=False)[0]


runAmiciSimulations = run_simulations
runAmiciSimulation = run_simulation
