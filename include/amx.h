/*
 * amx — C ABI of the MI355X (gfx950) batched stiff-ODE simulation and
 * sensitivity backend.  Every generated model module libamx_<model>.so exports
 * exactly these symbols.  All pointers in amx_batch_in / amx_batch_out and the
 * workspace are device pointers; nothing crosses this boundary as an
 * exception.
 *
 * Reference interfaces replaced:
 *   amx_run_batch      amici::run_simulations(Solver const&, std::vector<ExpData*> const&,
 *                        Model const&, bool failfast, int num_threads)
 *                        (include/amici/amici.h:36-39, src/amici.cpp:220-264) for
 *                        ODE models with sensitivity method none / forward / adjoint,
 *                        returning per-condition x, y, sx, llh, sllh and status
 *                        (ReturnData, include/amici/rdata.h:195-245).
 *   amx_solver_opts    the Solver options used on this path, with AMICI's defaults
 *                        (include/amici/solver.h:1869-2147).
 *   amx_model_info     model dimensions (ModelDimensions, include/amici/model_dimensions.h:59-156).
 *   amx_status_string  simulation_status_to_str_map (include/amici/amici.h).
 */
#ifndef AMX_H
#define AMX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMX_ABI_VERSION 10
#define AMX_NSTATS 18
#define AMX_NDIAG 9
/* bytes of the static workspace of amx_run_batch_static, per GPU */
#define AMX_STATIC_WORKSPACE_BYTES (64u << 20)

/* interpolation of the forward solution in the adjoint sweep
 * (amici::InterpolationType, include/amici/defines.h; CV_HERMITE / CV_POLYNOMIAL) */
#define AMX_INTERP_HERMITE 1
#define AMX_INTERP_POLYNOMIAL 2

/* BDF stability-limit detection (Solver::setStabilityLimitFlag, CVodeSetStabLimDet) */
#define AMX_STLDET_ON 1
#define AMX_STLDET_OFF 2

/* sensitivity method (amici::SensitivityMethod, include/amici/defines.h) */
#define AMX_SENSI_NONE 0
#define AMX_SENSI_FORWARD 1
#define AMX_SENSI_ADJOINT 2

typedef struct {
    double rtol, atol;         /* state tolerances (defaults 1e-8 / 1e-16) */
    double rtol_fsa, atol_fsa; /* forward sensitivities; NaN -> rtol / atol */
    double rtolB, atolB;       /* adjoint states; NaN -> rtol / atol */
    double quad_rtol, quad_atol; /* adjoint quadratures (defaults 1e-8 / 1e-12) */
    int maxsteps;              /* forward step limit (default 10000) */
    int maxstepsB;             /* backward step limit per interval; 0 -> 100 * maxsteps */
    int sensi_meth;            /* AMX_SENSI_* */
    int max_traj;              /* adjoint: stored forward steps per instance (0 -> maxsteps + 1) */
    double maxtime;            /* wall-clock seconds per instance and problem (0 = unlimited),
                                  AMICI_MAX_TIME_EXCEEDED when exceeded (Solver::setMaxTime) */
    int newton_maxsteps;       /* pre-equilibration: damped Newton iterations (Solver::setNewtonMaxSteps) */
    double ss_rtol, ss_atol;   /* steady-state tolerances; NaN -> 100 x rtol / atol
                                  (Solver ss_tol_factor, include/amici/solver.h) */
    int max_blocks;            /* 0: the persistent kernels get as many blocks as fit on the
                                  device at once; > 0: at most this many blocks, so fewer
                                  instances run at a time and each block takes further
                                  instances from the work queue (no counterpart in the
                                  reference: a launch-shape control)                    */
    int interp;                /* adjoint: AMX_INTERP_POLYNOMIAL (AMICI's default; 0 means
                                  the same) or AMX_INTERP_HERMITE (Solver::setInterpolationType) */
    int stldet;                /* AMX_STLDET_ON (AMICI's default, include/amici/solver.h:2037;
                                  0 means the same) or AMX_STLDET_OFF: BDF stability-limit
                                  detection in the forward and backward problems */
    int failfast;              /* nonzero: once an instance fails, instances with a larger
                                  index that have not started are skipped (status
                                  AMICI_NOT_RUN, llh / sllh NaN), as run_simulations' failfast
                                  (src/amici.cpp:232-260); which later instances had already
                                  started depends on timing, so a caller reports every
                                  instance after the first failed one as not run */
    int ckpt_steps;            /* adjoint: 0 (AMICI's setting, CVodeAdjInit(maxsteps + 1),
                                  src/solver_cvodes.cpp:726-739): every forward step is
                                  stored; K > 0: the forward pass writes a checkpoint of
                                  the integrator state every K steps and keeps only the
                                  last interval's steps, and the backward sweep
                                  recomputes each earlier interval from its checkpoint
                                  (CVodeAdjInit(K), CVodeF / CVAdataStore / CVodeB,
                                  cvodea.c:395-660, 1365-1630, 2328-2375).  The integrator
                                  takes a fresh Jacobian at every checkpoint (forward)
                                  and after every recomputation (backward), so the
                                  recomputed steps equal the forward pass's; max_traj
                                  then bounds the forward steps + 1 covered */
} amx_solver_opts;

typedef struct {
    int nb;                /* instances (conditions / parameter sets)                 */
    int nt;                /* output times per instance                               */
    int ts_stride;         /* 0: ts shared by all instances, nt: one row per instance */
    const double* theta;   /* [nb][np] free parameters on their scale (pscale; default log10) */
    const double* kfix;    /* [nb][nk] fixed parameters                               */
    const double* ts;      /* [1 or nb][nt] ascending, >= t0 = 0                      */
    const double* meas;    /* [nb][nt][ny] measurements, NaN = missing; may be NULL   */
    const double* sigma;   /* [nb][nt][ny] noise standard deviations                  */
    const double* kfix_preeq; /* [nb][nk] fixed parameters of a pre-equilibration to steady
                                 state (ExpData::fixed_parameters_pre_equilibration), or NULL;
                                 after it the fixed-parameter-dependent initial states are
                                 reset from kfix                                         */
    /* The following device arrays are optional (NULL: the default in brackets).
     * plist, nplist: sensitivities with respect to parameters plist[0..nplist-1]
     *   (Model::set_parameter_list, ExpData::plist; [all np in order]); sllh is
     *   then [nb][nplist] and sx [nb][nt][nplist][nx].
     * pscale: [np] ParameterScaling of each parameter, 0 none, 1 ln, 2 log10
     *   (Model::set_parameter_scale, ExpData::pscale; [all log10]); theta is on
     *   that scale and sllh / sx are derivatives with respect to it.
     *   A list or scale entry out of range fails every instance with
     *   AMICI_ILL_INPUT (-22).
     * x0: [nb][nx] initial states replacing the model's (ExpData::x0,
     *   Model::set_initial_state; [the model's]); a pre-equilibration starts
     *   from it too.
     * sx0: [nb][nplist][nx] their sensitivities with respect to the scaled
     *   parameters (ExpData::sx0, Model::set_initial_state_sensitivities;
     *   [0]); not together with kfix_preeq (-22).
     * t0: [nb] start time of each instance's simulation (ExpData t_start,
     *   Model::set_t0, applied in ConditionContext::apply_condition,
     *   src/edata.cpp:551; [0]); output times at or before it give the initial
     *   state, and the adjoint sweep ends there.                              */
    int nplist;
    const int* plist;
    const int* pscale;
    const double* x0;
    const double* sx0;
    const double* t0;
} amx_batch_in;

typedef struct {
    double* x;      /* [nb][nt][nx] or NULL                         */
    double* y;      /* [nb][nt][ny] or NULL                         */
    double* sx;     /* [nb][nt][nplist][nx] or NULL (forward only)  */
    double* llh;    /* [nb] log-likelihood (required)               */
    double* sllh;   /* [nb][nplist] gradient of llh w.r.t. the scaled parameters */
    int* status;    /* [nb] AMICI status code per instance (required) */
    int* stats;     /* [nb][AMX_NSTATS] solver counters or NULL      */
    int* diag;      /* [nb][nt][AMX_NDIAG] per-output-time diagnostics or NULL
                       (ReturnData numsteps, num_rhs_evals, num_err_test_fails,
                       num_non_lin_solv_conv_fails, order, numsteps_b, num_rhs_evals_b,
                       num_err_test_fails_b, num_non_lin_solv_conv_fails_b,
                       include/amici/rdata.h:261-301, filled as Solver::store_diagnosis /
                       store_diagnosis_b do, src/solver.cpp:349-398): slots 0-4 are
                       cumulative forward counters at output time it; slots 5-8 are
                       the backward problem's cumulative counters in the order the
                       backward sweep stores them (entry k at the k-th data point from
                       the last one), zero without an adjoint run */
} amx_batch_out;

/* dims[0..7] = nx, nr, np, nk, ny, team lanes per instance, noise model, ABI version */
int amx_model_info(int* dims);

/* Bytes of device workspace amx_run_batch needs for this batch / options. */
size_t amx_workspace_bytes(const amx_solver_opts* opts, int nb, int nt);

/* Runs the whole batch on `stream` (a hipStream_t; NULL = default stream).
 * Returns 0 when the launches were issued; per-instance results in
 * out->status after the stream completes.  Re-entrant: calls may run
 * concurrently from several threads or streams as long as each call in flight
 * has its own workspace (the work-queue counters and the adjoint trajectory
 * live there); calls sharing a workspace must be ordered on one stream.
 * Error codes:
 *   -22 invalid argument, -23 workspace too small, -30 HIP launch error.
 * Pre-equilibration failures are reported per instance as
 * AMICI_NO_STEADY_STATE (-81), AMICI_DAMPING_FACTOR_ERROR (-86) or
 * AMICI_SINGULAR_JACOBIAN (-9987) (include/amici/defines.h). */
int amx_run_batch(const amx_solver_opts* opts, const amx_batch_in* in, amx_batch_out* out,
                  void* workspace, size_t workspace_bytes, void* stream);

/* amx_run_batch without any allocation.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The workspace is a static
 * device array of AMX_STATIC_WORKSPACE_BYTES in the library, one per GPU,
 * held until the process ends; the caller still passes inputs and outputs.
 * The work-queue counters in it are cleared by the launch sequence itself.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one
 * GPU (other streams or threads) race: order them on one stream.  Returns the
 * codes of amx_run_batch, and -24 when the batch needs more workspace than
 * the array holds (amx_workspace_bytes() > amx_static_workspace_bytes(); a
 * smaller batch or opts->max_traj fits).  No hipMalloc / hipFree or
 * allocating library routine is called; the array's device address is looked
 * up once per GPU (hipGetSymbolAddress) on the first call made there. */
int amx_run_batch_static(const amx_solver_opts* opts, const amx_batch_in* in, amx_batch_out* out,
                         void* stream);
size_t amx_static_workspace_bytes(void);

/* Captured launch sequence (no counterpart in the reference: it replaces the
 * host loop over conditions of run_simulations, src/amici.cpp:235-262, by one
 * replayable HIP graph).  amx_graph_create records what amx_run_batch would
 * launch for these arguments (work-queue reset, forward kernel, backward
 * kernel) into a graph bound to these very buffers; amx_graph_launch replays
 * it on `stream`.  Inputs may be rewritten in place between replays; results
 * land in the same output buffers.  The buffers and the workspace must
 * outlive the graph.  Returns 0, or the error codes of amx_run_batch. */
typedef struct amx_graph amx_graph;
int amx_graph_create(const amx_solver_opts* opts, const amx_batch_in* in, amx_batch_out* out,
                     void* workspace, size_t workspace_bytes, amx_graph** graph);
int amx_graph_launch(amx_graph* graph, void* stream);
int amx_graph_destroy(amx_graph* graph);

/* Plain GPU BDF baseline: the same integrator, one instance per kernel launch
 * (sequential launches on `stream`), forward sensitivities or none. */
int amx_run_sequential(const amx_solver_opts* opts, const amx_batch_in* in, amx_batch_out* out,
                       void* workspace, size_t workspace_bytes, void* stream);

/* Diagnostics: record every accepted step and every failed attempt of the
 * forward integration of batch instance `instance` into the first half of
 * `buffer` and those of its backward (adjoint) integration into the second
 * half (device memory, `capacity` records of 8 doubles: event (1 step,
 * 2 nonlinear-solver failure, 3 error-test failure), t, h, order, error norm
 * or flag, cumulative Newton iterations, linear-solver setups, Jacobian
 * evaluations).  With adjoint checkpoints (ckpt_steps > 0) the backward
 * records take the third quarter of the buffer and the forward steps
 * recomputed from checkpoints the last quarter, interval by interval in the
 * order of recomputation (counters from 0 in each interval).  Applies to
 * the launches that follow, from any thread (a process-wide setting; the
 * library serialises access to it). */
int amx_set_trace(double* buffer, int instance, int capacity);

/* Benchmark instrumentation (no counterpart in the reference): while enabled,
 * every amx_run_batch / amx_run_sequential launch records HIP events on its
 * stream around its forward kernel and, for the adjoint, its backward kernel
 * (up to 4096 launches between reads).  Not for use under graph capture.
 * A process-wide setting; the library serialises access to the records. */
int amx_set_timing(int enable);

/* Waits for the recorded launches and returns the summed kernel durations
 * since the previous call: ms[0] forward kernels, ms[1] backward kernels,
 * launches[0] / launches[1] how many of each were timed.  Returns the number
 * of launches that were not recorded (table full), or -30 on a HIP error. */
int amx_kernel_times(double* ms, int* launches);

/* Text for an AMICI status code. */
const char* amx_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
