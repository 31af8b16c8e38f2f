/*
 * TEST INFRASTRUCTURE ONLY — the oracle (see amx_oracle.h).
 *
 * What this file restates (plain C, scalar) and where it comes from:
 *   - mass-action model functions w, xdot, J, dxdotdp     -> AMICI generated model
 *     semantics xdot = S*w (python/sdist/amici/importers/sbml/__init__.py:620-627),
 *     Model_ODE::fJ / fdxdotdp (src/model_ode.cpp:14-21, 141-170)
 *   - sensitivity RHS sxdot = J*sx + df/dp_ip             -> Model_ODE::fsxdot
 *     (src/model_ode.cpp:426-450; J and df/dp refreshed at ip == 0)
 *   - adjoint RHS xBdot = -J^T xB, qBdot = -(df/dp)^T xB  -> Model_ODE::fxBdot /
 *     fqBdot / fJB (src/model_ode.cpp:331-374)
 *   - Jy and dJy/dy of the continuous noise models         -> noise formulas
 *     (python/sdist/amici/importers/utils.py:434-461)
 *   - forward output loop with tstop = tout               -> EventHandlingSimulator::run
 *     (src/forwardproblem.cpp:53-151), is_next_t_too_close (:26-38),
 *     Solver::run (src/solver.cpp:141-160), apply_max_num_steps (:107-119)
 *   - llh / FSA sllh accumulation                          -> Model::add_observable_objective
 *     (_sensitivity) (src/model.cpp:1194-1239)
 *   - adjoint sweep with data-point jumps and the order-1 restart hack
 *                                                          -> EventHandlingBwdSimulator::run
 *     (src/backwardproblem.cpp:222-332), CVodeSolver::reset_state / reinit_b /
 *     reInit_post_process (src/solver_cvodes.cpp:471-649)
 *   - sllh = -(xB(t0).sx0 + xQB)                           -> ReturnData::process_backward_problem
 *     (src/rdata.cpp:514-629)
 *   - log10 chain rule on sllh / sx (non-zero entries only) -> rdata.cpp:749-866
 *   - solver options / tolerances                          -> Solver::setup / apply_tolerances*
 *     (src/solver.cpp:193-298, 631-691), defaults include/amici/solver.h:1869-2147
 * The integrator itself (BDF, Newton, dense LU, FSA, ASA, interpolation) is the
 * reference's own CVODES, compiled from ThirdParty/sundials by oracle/Makefile.
 *
 * Pinning (tests/test_oracle_*.py): closed-form linear ODE states, forward and
 * adjoint sensitivities; Robertson mass conservation and published reference
 * values; scipy Radau on the synthetic networks; central finite differences of
 * llh; ASA == FSA; replicate identity llh(2x data) = 2 llh; Jy closed forms.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <cvodes/cvodes.h>
#include <nvector/nvector_serial.h>
#include <sunlinsol/sunlinsol_dense.h>
#include <sunmatrix/sunmatrix_dense.h>

#include "cvodes_impl.h" /* internals touched by AMICI's backward reinit */

#include "amx_oracle.h"

#define AMICI_SUCCESS 0
#define AMICI_ERROR (-99)
#define AMICI_TOO_MUCH_WORK (-1)
#define AMICI_NO_STEADY_STATE (-81) /* include/amici/defines.h */

typedef struct {
    const amx_oracle_model* m;
    const double* p;
    double* w;      /* [nr] fluxes                  */
    double* mono;   /* [nr] w_r / p_r (monomials)    */
    double* J;      /* [nx*nx] row-major              */
    double* dfdp;   /* [nx*np] row-major              */
    double* u;      /* [nr] S^T v scratch            */
    const int* plist; /* sensitivity iS -> parameter plist[iS] (NULL: identity) */
    double* qfull;  /* [np] adjoint quadrature integrand of every parameter */
} ud_t;

/* ParameterScaling (include/amici/defines.h): 0 none, 1 ln, 2 log10;
 * unscale_parameters (src/misc.cpp) and the chain-rule factor of
 * ReturnData::apply_chain_rule_factor_to_simulation_results
 * (src/rdata.cpp:782-796: p ln 10, p, 1) */
static int pscale_of(const amx_oracle_opts* o, int j) { return o->pscale ? o->pscale[j] : 2; }
static double unscale(double v, int sc) { return sc == 2 ? pow(10.0, v) : (sc == 1 ? exp(v) : v); }
static double chain_coeff(double p, int sc) { return sc == 2 ? p * log(10.0) : (sc == 1 ? p : 1.0); }

/* ---------------- model pieces (mass action) ---------------- */

static void fluxes(const amx_oracle_model* m, const double* p, const double* x,
                   double* w, double* mono) {
    for (int r = 0; r < m->nr; ++r) {
        double prod = 1.0;
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) prod *= x[m->ridx[q]];
        mono[r] = prod;
        w[r] = p[m->kidx[r]] * prod;
    }
}

void amx_oracle_xdot(const amx_oracle_model* m, const double* p, const double* x,
                     double* xdot) {
    double* w = (double*)malloc(sizeof(double) * 2 * (size_t)m->nr);
    fluxes(m, p, x, w, w + m->nr);
    memset(xdot, 0, sizeof(double) * (size_t)m->nx);
    for (int r = 0; r < m->nr; ++r) {
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) xdot[m->ridx[q]] -= w[r];
        for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) xdot[m->pidx[q]] += w[r];
    }
    free(w);
}

/* d w_r / d x_s for every reactant position, scattered with the net stoichiometry. */
void amx_oracle_jac(const amx_oracle_model* m, const double* p, const double* x,
                    double* J) {
    const int nx = m->nx;
    memset(J, 0, sizeof(double) * (size_t)nx * nx);
    for (int r = 0; r < m->nr; ++r) {
        const double k = p[m->kidx[r]];
        for (int a = m->rptr[r]; a < m->rptr[r + 1]; ++a) {
            double d = k;
            for (int b = m->rptr[r]; b < m->rptr[r + 1]; ++b)
                if (b != a) d *= x[m->ridx[b]];
            const int s = m->ridx[a];
            for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) J[m->ridx[q] * nx + s] -= d;
            for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) J[m->pidx[q] * nx + s] += d;
        }
    }
}

void amx_oracle_dxdotdp(const amx_oracle_model* m, const double* p,
                        const double* x, double* dfdp) {
    (void)p;
    const int np = m->np;
    memset(dfdp, 0, sizeof(double) * (size_t)m->nx * np);
    for (int r = 0; r < m->nr; ++r) {
        double prod = 1.0;
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) prod *= x[m->ridx[q]];
        const int j = m->kidx[r];
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) dfdp[m->ridx[q] * np + j] -= prod;
        for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) dfdp[m->pidx[q] * np + j] += prod;
    }
}

/* The cost functions of AMICI's continuous noise distributions
 * (python/sdist/amici/importers/utils.py:434-461): noise 0 normal, 1 log-normal,
 * 2 log10-normal, 3 laplace, 4 log-laplace, 5 log10-laplace.  u(v) is the
 * model's scale (v, ln v or log10 v) and lnb its log base factor. */
static double noise_u(int sc, double v) { return sc == 0 ? v : (sc == 1 ? log(v) : log10(v)); }

double amx_oracle_Jy(int noise, double y, double sigma, double meas) {
    if (noise == 0) {
        const double r = (y - meas) / sigma;
        return 0.5 * log(2.0 * M_PI * sigma * sigma) + 0.5 * r * r;
    }
    if (noise == 1) {
        const double r = (log(y) - log(meas)) / sigma;
        return 0.5 * log(2.0 * M_PI * sigma * sigma * meas * meas) + 0.5 * r * r;
    }
    const int sc = noise % 3;
    const double r = (noise_u(sc, y) - noise_u(sc, meas)) / sigma;
    /* the measurement's density carries the Jacobian of the scale, m ln b */
    const double jac = sc == 0 ? 1.0 : (sc == 1 ? meas : meas * log(10.0));
    if (noise >= 3) return log(2.0 * sigma * jac) + fabs(r);
    return 0.5 * log(2.0 * M_PI * sigma * sigma * jac * jac) + 0.5 * r * r;
}

double amx_oracle_dJydy(int noise, double y, double sigma, double meas) {
    if (noise == 0) return (y - meas) / (sigma * sigma);
    if (noise == 1) return (log(y) - log(meas)) / (sigma * sigma * y);
    const int sc = noise % 3;
    const double r = (noise_u(sc, y) - noise_u(sc, meas)) / sigma;
    const double drdy = 1.0 / (sigma * (sc == 0 ? 1.0 : (sc == 1 ? y : y * log(10.0))));
    if (noise >= 3) return (r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : 0.0)) * drdy;
    return r * drdy;
}

/* ---------------- CVODES callbacks ---------------- */

static int cb_f(sunrealtype t, N_Vector x, N_Vector xdot, void* ud_) {
    (void)t;
    ud_t* ud = (ud_t*)ud_;
    const amx_oracle_model* m = ud->m;
    const double* xs = N_VGetArrayPointer(x);
    double* fd = N_VGetArrayPointer(xdot);
    fluxes(m, ud->p, xs, ud->w, ud->mono);
    memset(fd, 0, sizeof(double) * (size_t)m->nx);
    for (int r = 0; r < m->nr; ++r) {
        const double w = ud->w[r];
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) fd[m->ridx[q]] -= w;
        for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) fd[m->pidx[q]] += w;
    }
    for (int i = 0; i < m->nx; ++i)
        if (!isfinite(fd[i])) return 1; /* recoverable, as AMICI's check_finite */
    return 0;
}

static int cb_jac(sunrealtype t, N_Vector x, N_Vector fx, SUNMatrix Jm, void* ud_,
                  N_Vector t1, N_Vector t2, N_Vector t3) {
    (void)t; (void)fx; (void)t1; (void)t2; (void)t3;
    ud_t* ud = (ud_t*)ud_;
    const int nx = ud->m->nx;
    amx_oracle_jac(ud->m, ud->p, N_VGetArrayPointer(x), ud->J);
    for (int j = 0; j < nx; ++j) {
        double* col = SUNDenseMatrix_Column(Jm, j);
        for (int i = 0; i < nx; ++i) col[i] = ud->J[i * nx + j];
    }
    return 0;
}

static int cb_fS1(int Ns, sunrealtype t, N_Vector x, N_Vector xdot, int iS,
                  N_Vector sx, N_Vector sxdot, void* ud_, N_Vector t1, N_Vector t2) {
    (void)Ns; (void)t; (void)xdot; (void)t1; (void)t2;
    ud_t* ud = (ud_t*)ud_;
    const amx_oracle_model* m = ud->m;
    const int nx = m->nx, np = m->np;
    const double* xs = N_VGetArrayPointer(x);
    if (iS == 0) {
        amx_oracle_dxdotdp(m, ud->p, xs, ud->dfdp);
        amx_oracle_jac(m, ud->p, xs, ud->J);
    }
    const double* s = N_VGetArrayPointer(sx);
    double* o = N_VGetArrayPointer(sxdot);
    const int ip = ud->plist ? ud->plist[iS] : iS;
    for (int i = 0; i < nx; ++i) {
        double acc = ud->dfdp[i * np + ip];
        const double* Ji = ud->J + (size_t)i * nx;
        for (int j = 0; j < nx; ++j) acc += Ji[j] * s[j];
        o[i] = acc;
    }
    return 0;
}

/* xBdot = -J^T xB evaluated through the flux structure: u = S^T xB,
 * xBdot_s = -sum_r (dw_r/dx_s) u_r. Same arithmetic as the dense product. */
static int cb_fB(sunrealtype t, N_Vector x, N_Vector xB, N_Vector xBdot, void* ud_) {
    (void)t;
    ud_t* ud = (ud_t*)ud_;
    const amx_oracle_model* m = ud->m;
    const double* xs = N_VGetArrayPointer(x);
    const double* b = N_VGetArrayPointer(xB);
    double* o = N_VGetArrayPointer(xBdot);
    memset(o, 0, sizeof(double) * (size_t)m->nx);
    for (int r = 0; r < m->nr; ++r) {
        double u = 0.0;
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) u -= b[m->ridx[q]];
        for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) u += b[m->pidx[q]];
        const double k = ud->p[m->kidx[r]];
        for (int a = m->rptr[r]; a < m->rptr[r + 1]; ++a) {
            double d = k;
            for (int c = m->rptr[r]; c < m->rptr[r + 1]; ++c)
                if (c != a) d *= xs[m->ridx[c]];
            o[m->ridx[a]] -= d * u;
        }
    }
    return 0;
}

static int cb_fQB(sunrealtype t, N_Vector x, N_Vector xB, N_Vector qBdot, void* ud_) {
    (void)t;
    ud_t* ud = (ud_t*)ud_;
    const amx_oracle_model* m = ud->m;
    const double* xs = N_VGetArrayPointer(x);
    const double* b = N_VGetArrayPointer(xB);
    double* o = ud->qfull;
    memset(o, 0, sizeof(double) * (size_t)m->np);
    for (int r = 0; r < m->nr; ++r) {
        double u = 0.0;
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) u -= b[m->ridx[q]];
        for (int q = m->pptr[r]; q < m->pptr[r + 1]; ++q) u += b[m->pidx[q]];
        double prod = 1.0;
        for (int q = m->rptr[r]; q < m->rptr[r + 1]; ++q) prod *= xs[m->ridx[q]];
        o[m->kidx[r]] -= prod * u;
    }
    /* the backward problem integrates nplist quadratures (model.h nplist) */
    double* qd = N_VGetArrayPointer(qBdot);
    const int nq = (int)N_VGetLength(qBdot);
    for (int ip = 0; ip < nq; ++ip) qd[ip] = o[ud->plist ? ud->plist[ip] : ip];
    return 0;
}

static int cb_jacB(sunrealtype t, N_Vector x, N_Vector xB, N_Vector fB, SUNMatrix JB,
                   void* ud_, N_Vector t1, N_Vector t2, N_Vector t3) {
    (void)t; (void)xB; (void)fB; (void)t1; (void)t2; (void)t3;
    ud_t* ud = (ud_t*)ud_;
    const int nx = ud->m->nx;
    amx_oracle_jac(ud->m, ud->p, N_VGetArrayPointer(x), ud->J);
    for (int j = 0; j < nx; ++j) {
        double* col = SUNDenseMatrix_Column(JB, j);
        for (int i = 0; i < nx; ++i) col[i] = -ud->J[j * nx + i];
    }
    return 0;
}

/* ---------------- AMICI layer restatement ---------------- */

/* forwardproblem.cpp:26-38 */
static int next_t_too_close(double cur, double next) {
    const double tdiff = next - cur;
    if (tdiff == 0.0) return 1;
    const double tround = 2.220446049250313e-16 * fmax(fabs(cur), fabs(next));
    return fabs(tdiff) < 2.0 * tround;
}

/* solver_cvodes.cpp:471-505: force order 1 and overwrite zn[0]. */
static void reset_state(CVodeMem cv_mem, N_Vector y0) {
    cv_mem->cv_q = 1;
    cv_mem->cv_L = 2;
    cv_mem->cv_qwait = cv_mem->cv_L;
    cv_mem->cv_qu = 0;
    cv_mem->cv_hu = 0.0;
    cv_mem->cv_tolsf = 1.0;
    cv_mem->cv_h0u = 0.0;
    cv_mem->cv_next_h = 0.0;
    cv_mem->cv_next_q = 0;
    N_VScale(1.0, y0, cv_mem->cv_zn[0]);
}

/* solver_cvodes.cpp:527-551 for a backward problem: one ONE_STEP call with nst
 * temporarily zeroed so CVode re-runs its first-call initialisation. */
static int reinit_post_process_b(void* cvode_mem, N_Vector xB, double tnext) {
    CVodeMem cv_mem = (CVodeMem)cvode_mem;
    CVadjMem ca_mem = cv_mem->cv_adj_mem;
    CVodeBMem cvB_mem = ca_mem->cvB_mem;
    while (cvB_mem != NULL) {
        ca_mem->ca_bckpbCrt = cvB_mem;
        CVodeMem bmem = cvB_mem->cv_mem;
        long nst_tmp = bmem->cv_nst;
        bmem->cv_nst = 0;
        if (CVodeSetStopTime(bmem, tnext) != CV_SUCCESS) return AMICI_ERROR;
        double tret;
        int status = CVode(bmem, tnext, xB, &tret, CV_ONE_STEP);
        if (status != CV_SUCCESS) return status < 0 ? status : AMICI_ERROR;
        bmem->cv_nst = nst_tmp + 1;
        cvB_mem->cv_tout = tret;
        cvB_mem = cvB_mem->cv_next;
    }
    return AMICI_SUCCESS;
}

#define CHECK(call)                                  \
    do {                                             \
        int _s = (call);                             \
        if (_s < 0) { status = _s; goto cleanup; }   \
    } while (0)

/* Pre-equilibration (SteadyStateProblem, src/forwardproblem.cpp:755-1223,
 * restated by integration instead of Newton's method): CVODES from
 * x0(kpre) to t = o->t_preeq, with forward sensitivities from sx0 = 0 when
 * sx != NULL; the state there must satisfy AMICI's steady-state criterion
 * WRMS(xdot) < 1 with tolerances 100 x (rtol, atol) (solver.h ss_tol_factor).
 * On return x holds x_ss and sx (if given) dx_ss/dp (linear parameters). */
static int preequilibrate(const amx_oracle_model* m, const amx_oracle_opts* o, ud_t* ud,
                          const double* kpre, int x0_given, N_Vector x, N_Vector* sx,
                          SUNContext ctx) {
    const int nx = m->nx, np = m->np;
    int status = AMICI_SUCCESS;
    void* cv = NULL;
    SUNMatrix A = NULL;
    SUNLinearSolver LS = NULL;
    int* plist = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    double* atols = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
    for (int j = 0; j < np; ++j) { plist[j] = j; atols[j] = o->atol; }
    const int* plist_saved = ud->plist;
    ud->plist = NULL;
    double* xd = N_VGetArrayPointer(x);
    /* x holds the initial state on entry; without ExpData::x0 the
     * fixed-parameter-dependent species start from kpre (Model::initialize_states) */
    if (!x0_given)
        for (int i = 0; i < nx; ++i)
            if (m->x0k[i] >= 0) xd[i] = kpre[m->x0k[i]];
    cv = CVodeCreate(CV_BDF, ctx);
    CHECK(CVodeInit(cv, cb_f, 0.0, x));
    CHECK(CVodeSStolerances(cv, o->rtol, o->atol));
    CHECK(CVodeSetUserData(cv, ud));
    A = SUNDenseMatrix(nx, nx, ctx);
    LS = SUNLinSol_Dense(x, A, ctx);
    CHECK(CVodeSetLinearSolver(cv, LS, A));
    CHECK(CVodeSetJacFn(cv, cb_jac));
    CHECK(CVodeSetMaxNumSteps(cv, o->maxsteps));
    if (sx) {
        for (int j = 0; j < np; ++j) N_VConst(0.0, sx[j]);
        CHECK(CVodeSensInit1(cv, np, o->ism, cb_fS1, sx));
        CHECK(CVodeSetSensParams(cv, (double*)ud->p, NULL, plist));
        CHECK(CVodeSensSStolerances(cv, o->rtol, atols));
        CHECK(CVodeSetSensErrCon(cv, SUNTRUE));
    }
    {
        double t = 0.0;
        int rc = CVode(cv, o->t_preeq, x, &t, CV_NORMAL);
        if (rc < 0) { status = rc; goto cleanup; }
        if (sx) CHECK(CVodeGetSens(cv, &t, sx));
        double* f = (double*)malloc(sizeof(double) * (size_t)nx);
        amx_oracle_xdot(m, ud->p, xd, f);
        double acc = 0.0;
        for (int i = 0; i < nx; ++i) {
            const double e = f[i] / (100.0 * o->rtol * fabs(xd[i]) + 100.0 * o->atol);
            acc += e * e;
        }
        free(f);
        if (!(sqrt(acc / nx) < 1.0)) status = AMICI_NO_STEADY_STATE;
    }
cleanup:
    ud->plist = plist_saved;
    if (cv) CVodeFree(&cv);
    if (LS) SUNLinSolFree(LS);
    if (A) SUNMatDestroy(A);
    free(plist); free(atols);
    return status;
}

int amx_oracle_simulate(const amx_oracle_model* m, const amx_oracle_opts* o,
                        const double* theta, const double* kfix, const double* kpre,
                        const double* x0, const double* sx0_in, int nt,
                        const double* ts, const double* meas, const double* sigma,
                        double* x_out, double* y_out, double* sx_out, double* llh_out,
                        double* sllh_out, long* counters) {
    const int nx = m->nx, ny = m->ny, np_all = m->np;
    /* sensitivities w.r.t. plist (edata.cpp:446-447, Model::set_parameter_list) */
    const int np = o->plist ? o->nplist : m->np;
    const double t0 = 0.0;
    int status = AMICI_SUCCESS;
    SUNContext ctx = NULL;
    void* cv = NULL;
    N_Vector x = NULL, xB = NULL, qB = NULL;
    N_Vector* sx = NULL;
    SUNMatrix A = NULL, AB = NULL;
    SUNLinearSolver LS = NULL, LSB = NULL;
    int which = 0;
    double llh = 0.0;

    double* p = (double*)malloc(sizeof(double) * (size_t)(np_all > 0 ? np_all : 1));
    double* work = (double*)calloc((size_t)(3 * m->nr + nx * nx + nx * np_all + nx + np_all),
                                   sizeof(double));
    double* dJydx = (double*)calloc((size_t)nt * nx, sizeof(double));
    double* sllh = (double*)calloc((size_t)(np > 0 ? np : 1), sizeof(double));
    double* ybuf = (double*)malloc(sizeof(double) * (size_t)(ny > 0 ? ny : 1));
    /* [np_all][nx] initial state sensitivities of every parameter */
    double* sx0 = (double*)calloc((size_t)nx * (np_all > 0 ? np_all : 1), sizeof(double));
    int* plist = (int*)malloc(sizeof(int) * (size_t)(np > 0 ? np : 1));
    double* atols = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
    ud_t ud;
    ud.m = m;
    ud.p = p;
    ud.w = work;
    ud.mono = work + m->nr;
    ud.u = work + 2 * m->nr;
    ud.J = work + 3 * m->nr;
    ud.dfdp = ud.J + nx * nx;
    ud.qfull = ud.dfdp + nx * np_all + nx;
    ud.plist = plist;
    if (counters) memset(counters, 0, sizeof(long) * AMX_ORACLE_NCOUNTERS);

    /* edata.cpp:438-557 + model.cpp:559: p unscaled per parameter scale */
    for (int j = 0; j < np_all; ++j) p[j] = unscale(theta[j], pscale_of(o, j));
    for (int j = 0; j < np; ++j) { plist[j] = o->plist ? o->plist[j] : j; atols[j] = o->atol; }

    if (SUNContext_Create(SUN_COMM_NULL, &ctx) != 0) { status = AMICI_ERROR; goto cleanup; }
    x = N_VNew_Serial(nx, ctx);
    {
        /* initial state: ExpData::x0 if given (Model::set_initial_state),
         * else the model's x0 (constants or fixed parameters) */
        double* xd = N_VGetArrayPointer(x);
        for (int i = 0; i < nx; ++i)
            xd[i] = x0 ? x0[i] : ((m->x0k[i] >= 0) ? kfix[m->x0k[i]] : m->x0c[i]);
    }
    if (kpre) {
        /* pre-equilibration, then the fixed-parameter-dependent initial states
         * are re-initialised from kfix (reinitialize_fixed_parameter_initial_states) */
        N_Vector* sxp = (o->sensi >= 1 && np > 0) ? N_VCloneVectorArray(np_all, x) : NULL;
        status = preequilibrate(m, o, &ud, kpre, x0 != NULL, x, sxp, ctx);
        if (sxp) {
            for (int j = 0; j < np_all; ++j) {
                const double* s = N_VGetArrayPointer(sxp[j]);
                for (int i = 0; i < nx; ++i) sx0[(size_t)j * nx + i] = (m->x0k[i] >= 0) ? 0.0 : s[i];
            }
            N_VDestroyVectorArray(sxp, np_all);
        }
        if (status != AMICI_SUCCESS) goto cleanup;
        double* xd = N_VGetArrayPointer(x);
        for (int i = 0; i < nx; ++i)
            if (m->x0k[i] >= 0) xd[i] = kfix[m->x0k[i]];
    } else if (sx0_in) {
        /* ExpData::sx0 (Model::set_initial_state_sensitivities, model.cpp:986-1023):
         * given w.r.t. the scaled parameters, stored divided by the chain-rule factor */
        for (int j = 0; j < np; ++j) {
            const double cf = chain_coeff(p[plist[j]], pscale_of(o, plist[j]));
            for (int i = 0; i < nx; ++i)
                sx0[(size_t)plist[j] * nx + i] = sx0_in[(size_t)j * nx + i] / cf;
        }
    }
    cv = CVodeCreate(CV_BDF, ctx);
    CHECK(CVodeInit(cv, cb_f, t0, x));
    CHECK(CVodeSStolerances(cv, o->rtol, o->atol));
    CHECK(CVodeSetUserData(cv, &ud));
    CHECK(CVodeSetStabLimDet(cv, o->stldet ? SUNTRUE : SUNFALSE));
    A = SUNDenseMatrix(nx, nx, ctx);
    LS = SUNLinSol_Dense(x, A, ctx);
    CHECK(CVodeSetLinearSolver(cv, LS, A));
    CHECK(CVodeSetJacFn(cv, cb_jac));
    CHECK(CVodeSetMaxNonlinIters(cv, 3));
    CHECK(CVodeSetMaxConvFails(cv, 10));
    if (o->sensi == 1 && np > 0) {
        sx = N_VCloneVectorArray(np, x);
        for (int j = 0; j < np; ++j) /* sx0 = dx0/dp: 0, or from pre-equilibration */
            memcpy(N_VGetArrayPointer(sx[j]), sx0 + (size_t)plist[j] * nx, sizeof(double) * (size_t)nx);
        CHECK(CVodeSensInit1(cv, np, o->ism, cb_fS1, sx));
        CHECK(CVodeSetSensParams(cv, p, NULL, plist));
        CHECK(CVodeSensSStolerances(cv, o->rtol, atols));
        CHECK(CVodeSetSensErrCon(cv, SUNTRUE));
    } else if (o->sensi == 2 && np > 0) {
        CHECK(CVodeAdjInit(cv, o->maxsteps + 1, o->interp));
    }

    /* ---- forward: EventHandlingSimulator::run ---- */
    double t = t0;
    int ncheck = 0;
    for (int it = 0; it < nt; ++it) {
        const double tout = ts[it];
        if (tout > t0) {
            while (t < tout) {
                if (next_t_too_close(t, tout)) break;
                CHECK(CVodeSetStopTime(cv, tout));
                long cur = 0;
                CVodeGetNumSteps(cv, &cur);
                /* budget used up at an output time: AmiException in
                   Solver::apply_max_num_steps (src/solver.cpp:107-119) -> AMICI_ERROR */
                if (o->maxsteps <= cur) { status = AMICI_ERROR; goto cleanup; }
                CHECK(CVodeSetMaxNumSteps(cv, o->maxsteps - cur));
                int s = (o->sensi == 2 && np > 0)
                            ? CVodeF(cv, tout, x, &t, CV_NORMAL, &ncheck)
                            : CVode(cv, tout, x, &t, CV_NORMAL);
                if (s < 0) { status = s; goto cleanup; }
            }
            if (sx) { double td; CHECK(CVodeGetSens(cv, &td, sx)); }
        }
        const double* xd = N_VGetArrayPointer(x);
        if (x_out) memcpy(x_out + (size_t)it * nx, xd, sizeof(double) * (size_t)nx);
        if (sx_out && sx)
            for (int j = 0; j < np; ++j)
                memcpy(sx_out + ((size_t)it * np + j) * nx, N_VGetArrayPointer(sx[j]),
                       sizeof(double) * (size_t)nx);
        /* observables, Jy, dJydx (model.cpp:1194-1259, 2382-2413) */
        for (int iy = 0; iy < ny; ++iy) {
            double acc = 0.0;
            for (int i = 0; i < nx; ++i) acc += m->C[iy * nx + i] * xd[i];
            ybuf[iy] = acc;
            if (y_out) y_out[(size_t)it * ny + iy] = acc;
        }
        for (int iy = 0; iy < ny; ++iy) {
            const double mv = meas ? meas[(size_t)it * ny + iy] : NAN;
            if (isnan(mv)) continue;
            const double sg = sigma[(size_t)it * ny + iy];
            llh -= amx_oracle_Jy(m->noise, ybuf[iy], sg, mv);
            const double g = amx_oracle_dJydy(m->noise, ybuf[iy], sg, mv);
            for (int i = 0; i < nx; ++i) dJydx[(size_t)it * nx + i] += g * m->C[iy * nx + i];
        }
        if (sx) /* sllh -= dJydx . sx */
            for (int j = 0; j < np; ++j) {
                const double* s = N_VGetArrayPointer(sx[j]);
                double acc = 0.0;
                for (int i = 0; i < nx; ++i) acc += dJydx[(size_t)it * nx + i] * s[i];
                sllh[j] -= acc;
            }
    }
    if (counters) {
        long v;
        CVodeGetNumSteps(cv, &v); counters[0] = v;
        CVodeGetNumRhsEvals(cv, &v); counters[1] = v;
        CVodeGetNumLinSolvSetups(cv, &v); counters[2] = v;
        CVodeGetNumJacEvals(cv, &v); counters[3] = v;
        CVodeGetNumNonlinSolvIters(cv, &v); counters[4] = v;
        CVodeGetNumErrTestFails(cv, &v); counters[5] = v;
        CVodeGetNumNonlinSolvConvFails(cv, &v); counters[6] = v;
        CVodeGetNumStabLimOrderReds(cv, &v); counters[15] = v;
        if (sx) {
            CVodeGetSensNumRhsEvals(cv, &v); counters[7] = v;
            if (o->ism != CV_SIMULTANEOUS) { CVodeGetSensNumNonlinSolvIters(cv, &v); counters[8] = v; }
        }
    }

    /* ---- backward: EventHandlingBwdSimulator::run ---- */
    if (o->sensi == 2 && np > 0 && nt > 0) {
        int it = nt - 1;
        double tcur = (ts[it] > t0) ? t : t0; /* forward final time */
        xB = N_VNew_Serial(nx, ctx);
        qB = N_VNew_Serial(np, ctx);
        N_VConst(0.0, xB);
        N_VConst(0.0, qB);
        double* b = N_VGetArrayPointer(xB);
        if (ts[it] == tcur) {
            for (int i = 0; i < nx; ++i) b[i] += dJydx[(size_t)it * nx + i];
            --it;
        }
        if (tcur > t0 || it >= 0) {
            CHECK(CVodeCreateB(cv, CV_BDF, &which));
            CHECK(CVodeInitB(cv, which, cb_fB, tcur, xB));
            CHECK(CVodeSetUserDataB(cv, which, &ud));
            AB = SUNDenseMatrix(nx, nx, ctx);
            LSB = SUNLinSol_Dense(xB, AB, ctx);
            CHECK(CVodeSetLinearSolverB(cv, which, LSB, AB));
            CHECK(CVodeSetJacFnB(cv, which, cb_jacB));
            CHECK(CVodeQuadInitB(cv, which, cb_fQB, qB));
            CHECK(CVodeSStolerancesB(cv, which, o->rtol, o->atol));
            CHECK(CVodeSetQuadErrConB(cv, which,
                                      (isfinite(o->quad_atol) && isfinite(o->quad_rtol)) ? SUNTRUE : SUNFALSE));
            CHECK(CVodeQuadSStolerancesB(cv, which, o->quad_rtol, o->quad_atol));
            CHECK(CVodeSetStabLimDetB(cv, which, o->stldet ? SUNTRUE : SUNFALSE));
            CHECK(CVodeSetMaxNumStepsB(cv, which, o->maxsteps * 100));
            int called = 0, force_pp = 0;
            void* cvB = CVodeGetAdjCVodeBmem(cv, which);
            while (it >= 0) {
                const double tnext = ts[it];
                if (tnext < tcur) {
                    if (force_pp) {
                        int s = reinit_post_process_b(cv, xB, tnext);
                        if (s != AMICI_SUCCESS) { status = s; goto cleanup; }
                        force_pp = 0;
                    }
                    int s = CVodeB(cv, tnext, CV_NORMAL);
                    if (s != CV_SUCCESS && s != CV_TSTOP_RETURN) { status = s < 0 ? s : AMICI_ERROR; goto cleanup; }
                    called = 1;
                    double tr;
                    CHECK(CVodeGetB(cv, which, &tr, xB));
                    CHECK(CVodeGetQuadB(cv, which, &tr, qB));
                    tcur = tr;
                }
                for (int i = 0; i < nx; ++i) b[i] += dJydx[(size_t)it * nx + i];
                --it;
                /* reinit_b + reinit_quad_b (solver_cvodes.cpp:594-649) */
                CVodeMem bm = (CVodeMem)cvB;
                if (called) force_pp = 1;
                bm->cv_tn = tcur;
                reset_state(bm, xB);
                N_VScale(1.0, qB, bm->cv_znQ[0]);
            }
            if (tcur > t0) {
                if (force_pp) {
                    int s = reinit_post_process_b(cv, xB, t0);
                    if (s != AMICI_SUCCESS) { status = s; goto cleanup; }
                }
                int s = CVodeB(cv, t0, CV_NORMAL);
                if (s != CV_SUCCESS && s != CV_TSTOP_RETURN) { status = s < 0 ? s : AMICI_ERROR; goto cleanup; }
                double tr;
                CHECK(CVodeGetB(cv, which, &tr, xB));
                CHECK(CVodeGetQuadB(cv, which, &tr, qB));
            }
            if (counters) {
                long v;
                CVodeGetNumSteps(cvB, &v); counters[9] = v;
                CVodeGetNumRhsEvals(cvB, &v); counters[10] = v;
                CVodeGetNumLinSolvSetups(cvB, &v); counters[11] = v;
                CVodeGetNumNonlinSolvIters(cvB, &v); counters[12] = v;
                CVodeGetQuadNumRhsEvals(cvB, &v); counters[13] = v;
                CVodeGetNumJacEvals(cvB, &v); counters[14] = v;
            }
        }
        /* rdata.cpp:514-629: sllh = -(xB(t0) . sx0 + xQB); sx0 = 0 unless the
         * instance was pre-equilibrated */
        const double* qd = N_VGetArrayPointer(qB);
        const double* bd = N_VGetArrayPointer(xB);
        for (int j = 0; j < np; ++j) {
            double xs = 0.0;
            for (int i = 0; i < nx; ++i) xs += bd[i] * sx0[(size_t)plist[j] * nx + i];
            sllh[j] -= xs + qd[j];
        }
    }

    /* rdata.cpp:782-830: chain rule of each parameter's scale on non-zero entries */
    if (o->sensi >= 1) {
        for (int j = 0; j < np; ++j)
            if (sllh[j] != 0.0) sllh[j] *= chain_coeff(p[plist[j]], pscale_of(o, plist[j]));
        if (sx_out && sx)
            for (int it = 0; it < nt; ++it)
                for (int j = 0; j < np; ++j)
                    for (int i = 0; i < nx; ++i) {
                        double* v = sx_out + ((size_t)it * np + j) * nx + i;
                        if (*v != 0.0) *v *= chain_coeff(p[plist[j]], pscale_of(o, plist[j]));
                    }
    }
    if (llh_out) *llh_out = llh;
    if (sllh_out && o->sensi >= 1) memcpy(sllh_out, sllh, sizeof(double) * (size_t)np);

cleanup:
    if (status != AMICI_SUCCESS && llh_out) *llh_out = NAN;
    if (sx) N_VDestroyVectorArray(sx, np);
    if (cv) CVodeFree(&cv);
    if (LS) SUNLinSolFree(LS);
    if (LSB) SUNLinSolFree(LSB);
    if (A) SUNMatDestroy(A);
    if (AB) SUNMatDestroy(AB);
    if (x) N_VDestroy(x);
    if (xB) N_VDestroy(xB);
    if (qB) N_VDestroy(qB);
    if (ctx) SUNContext_Free(&ctx);
    free(p); free(work); free(dJydx); free(sllh); free(ybuf); free(plist); free(atols); free(sx0);
    return status;
}

int amx_oracle_simulate_batch(const amx_oracle_model* m, const amx_oracle_opts* o,
                              int nb, const double* theta, const double* kfix,
                              const double* kpre, const double* x0, const double* sx0,
                              int nt, const double* ts,
                              const double* meas,
                              const double* sigma, double* x, double* y,
                              double* sx, double* llh, double* sllh,
                              long* counters, int* status, int nthreads) {
    const size_t nx = (size_t)m->nx, ny = (size_t)m->ny, np = (size_t)m->np, nk = (size_t)m->nk;
    const size_t npl = o->plist ? (size_t)o->nplist : np;
    int nfail = 0;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1) reduction(+ : nfail)
    for (int b = 0; b < nb; ++b) {
        const size_t B = (size_t)b;
        int s = amx_oracle_simulate(
            m, o, theta + B * np, kfix + B * nk, kpre ? kpre + B * nk : NULL,
            x0 ? x0 + B * nx : NULL, sx0 ? sx0 + B * npl * nx : NULL, nt, ts + B * (size_t)nt,
            meas ? meas + B * (size_t)nt * ny : NULL, sigma + B * (size_t)nt * ny,
            x ? x + B * (size_t)nt * nx : NULL, y ? y + B * (size_t)nt * ny : NULL,
            sx ? sx + B * (size_t)nt * npl * nx : NULL, llh ? llh + B : NULL,
            sllh ? sllh + B * npl : NULL,
            counters ? counters + B * AMX_ORACLE_NCOUNTERS : NULL);
        if (status) status[b] = s;
        if (s != 0) nfail++;
    }
    return nfail;
}

/* Diagnostics (test infrastructure): the forward integration of one instance
 * (states only, x0 from kfix, tstop at every output time as in
 * amx_oracle_simulate) driven in CV_ONE_STEP mode, which takes the same steps
 * as CV_NORMAL mode and returns after each.  Every accepted step is recorded
 * as 9 doubles: t, step size used, order used, nst, nfe, nsetups, nje, netf,
 * ncfn.  Returns the number of records (at most cap) or a negative status. */
void (*amx_oracle_step_hook)(void* cvode_mem) = NULL;

int amx_oracle_trace_forward(const amx_oracle_model* m, const amx_oracle_opts* o,
                             const double* theta, const double* kfix, int nt, const double* ts,
                             double* rec, int cap) {
    const int nx = m->nx, np = m->np;
    int status = 0, n = 0;
    SUNContext ctx = NULL;
    void* cv = NULL;
    N_Vector x = NULL;
    SUNMatrix A = NULL;
    SUNLinearSolver LS = NULL;
    double* p = (double*)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
    double* work = (double*)calloc((size_t)(3 * m->nr + nx * nx + nx * np + nx), sizeof(double));
    ud_t ud;
    ud.m = m;
    ud.p = p;
    ud.w = work;
    ud.mono = work + m->nr;
    ud.u = work + 2 * m->nr;
    ud.J = work + 3 * m->nr;
    ud.dfdp = ud.J + nx * nx;
    ud.plist = NULL;
    ud.qfull = NULL;
    for (int j = 0; j < np; ++j) p[j] = unscale(theta[j], pscale_of(o, j));
    if (SUNContext_Create(SUN_COMM_NULL, &ctx) != 0) { status = AMICI_ERROR; goto cleanup; }
    x = N_VNew_Serial(nx, ctx);
    {
        double* xd = N_VGetArrayPointer(x);
        for (int i = 0; i < nx; ++i) xd[i] = (m->x0k[i] >= 0) ? kfix[m->x0k[i]] : m->x0c[i];
    }
    cv = CVodeCreate(CV_BDF, ctx);
    CHECK(CVodeInit(cv, cb_f, 0.0, x));
    CHECK(CVodeSStolerances(cv, o->rtol, o->atol));
    CHECK(CVodeSetUserData(cv, &ud));
    CHECK(CVodeSetStabLimDet(cv, o->stldet ? SUNTRUE : SUNFALSE));
    A = SUNDenseMatrix(nx, nx, ctx);
    LS = SUNLinSol_Dense(x, A, ctx);
    CHECK(CVodeSetLinearSolver(cv, LS, A));
    CHECK(CVodeSetJacFn(cv, cb_jac));
    CHECK(CVodeSetMaxNonlinIters(cv, 3));
    CHECK(CVodeSetMaxConvFails(cv, 10));
    CHECK(CVodeSetMaxNumSteps(cv, o->maxsteps));
    {
        double t = 0.0;
        for (int it = 0; it < nt; ++it) {
            const double tout = ts[it];
            if (tout <= 0.0) continue;
            while (t < tout) {
                if (next_t_too_close(t, tout)) break;
                CHECK(CVodeSetStopTime(cv, tout));
                const int s = CVode(cv, tout, x, &t, CV_ONE_STEP);
                if (s < 0) { status = s; goto cleanup; }
                if (n < cap) {
                    long v;
                    double hu;
                    int qu;
                    double* r = rec + (size_t)n * 9;
                    CVodeGetCurrentTime(cv, &r[0]);
                    CVodeGetLastStep(cv, &hu);
                    CVodeGetLastOrder(cv, &qu);
                    r[1] = hu;
                    r[2] = qu;
                    CVodeGetNumSteps(cv, &v); r[3] = (double)v;
                    CVodeGetNumRhsEvals(cv, &v); r[4] = (double)v;
                    CVodeGetNumLinSolvSetups(cv, &v); r[5] = (double)v;
                    CVodeGetNumJacEvals(cv, &v); r[6] = (double)v;
                    CVodeGetNumErrTestFails(cv, &v); r[7] = (double)v;
                    CVodeGetNumNonlinSolvConvFails(cv, &v); r[8] = (double)v;
                }
                if (amx_oracle_step_hook) amx_oracle_step_hook(cv);
                ++n;
            }
        }
    }
cleanup:
    if (cv) CVodeFree(&cv);
    if (LS) SUNLinSolFree(LS);
    if (A) SUNMatDestroy(A);
    if (x) N_VDestroy(x);
    if (ctx) SUNContext_Free(&ctx);
    free(p);
    free(work);
    return status < 0 ? status : n;
}
