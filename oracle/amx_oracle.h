/*
 * TEST INFRASTRUCTURE ONLY — the oracle.
 *
 * CPU restatement of AMICI's model / objective / forward- and backward-problem
 * layer for mass-action reaction networks, driven on top of the real SUNDIALS
 * CVODES 7.1.1 that ships with the reference (built from its sources into
 * oracle/_ref/ by oracle/Makefile).  Only the tests, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker / reported CPU baseline; the product never calls it.
 *
 * Parity pinning of each restated function is listed in the header of
 * cvodes_harness.c and checked by tests/test_oracle_*.py.
 */
#ifndef AMX_ORACLE_H
#define AMX_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Mass-action reaction network (flattened tables).
 *   rate of reaction r:  w_r = p[kidx[r]] * prod_{m in reactants(r)} x[m]
 *   xdot = S w,  S[i][r] = (#i in products(r)) - (#i in reactants(r))
 * Mirrors AMICI's SBML mass-action import xdot = S*w
 * (python/sdist/amici/importers/sbml/__init__.py:620-627). */
typedef struct {
    int nx;            /* species / states                              */
    int nr;            /* reactions                                     */
    int np;            /* free parameters (rate constants), log10-scaled */
    int nk;            /* fixed parameters (initial concentrations)      */
    int ny;            /* observables                                   */
    const int* rptr;   /* [nr+1] reactant CSR (multiplicity by repeat)  */
    const int* ridx;
    const int* pptr;   /* [nr+1] product CSR                            */
    const int* pidx;
    const int* kidx;   /* [nr] index of the rate constant in p          */
    const double* x0c; /* [nx] constant initial concentration           */
    const int* x0k;    /* [nx] -1, or index into k giving x0            */
    const double* C;   /* [ny*nx] row-major observable matrix y = C x   */
    int noise;         /* 0 normal, 1 log-normal, 2 log10-normal, 3 laplace,
                          4 log-laplace, 5 log10-laplace */
} amx_oracle_model;

typedef struct {
    double rtol, atol;       /* state tolerances (also FSA/ASA, as AMICI's defaults) */
    double quad_rtol, quad_atol;
    long maxsteps;
    int sensi;               /* 0 none, 1 forward (FSA), 2 adjoint (ASA) */
    int ism;                 /* CV_SIMULTANEOUS=1, CV_STAGGERED=2         */
    int interp;              /* CV_HERMITE=1, CV_POLYNOMIAL=2             */
    int stldet;              /* stability limit detection (AMICI default on) */
    double t_preeq;          /* pre-equilibration: integration end time */
    int nplist;              /* sensitivities w.r.t. plist[0..nplist-1] (Model::set_parameter_list);
                                plist NULL: all np parameters in order */
    const int* plist;
    const int* pscale;       /* [np] ParameterScaling (0 none, 1 ln, 2 log10) of theta, or NULL:
                                all log10 (Model::set_parameter_scale) */
} amx_oracle_opts;

/* Counters: [0]nst [1]nfe [2]nsetups [3]nje [4]nni [5]netf [6]ncfn
 *           [7]nfSe [8]nniS [9]nstB [10]nfeB [11]nsetupsB [12]nniB [13]nfQeB [14]njeB
 *           [15]nor (forward order reductions by stability-limit detection) */
#define AMX_ORACLE_NCOUNTERS 16

/* One condition.  theta: [np] parameters on their scale (pscale); kfix: [nk];
 * kpre: [nk] fixed parameters of a pre-equilibration (NULL = none); x0: [nx]
 * initial state replacing the model's (ExpData::x0, NULL = none); sx0:
 * [nplist*nx] initial state sensitivities with respect to the scaled
 * parameters (ExpData::sx0, NULL = none; not with kpre); ts: [nt];
 * meas/sigma: [nt*ny] (NaN measurement = missing).
 * Outputs (may be NULL): x [nt*nx], y [nt*ny], sx [nt*nplist*nx] (FSA only),
 * llh [1], sllh [nplist] (w.r.t. the scaled parameters), counters
 * [AMX_ORACLE_NCOUNTERS].
 * Returns AMICI status (0 success, <0 failure as in include/amici/defines.h). */
int amx_oracle_simulate(const amx_oracle_model* m, const amx_oracle_opts* o,
                        const double* theta, const double* kfix, const double* kpre,
                        const double* x0, const double* sx0, int nt,
                        const double* ts, const double* meas, const double* sigma,
                        double* x, double* y, double* sx, double* llh,
                        double* sllh, long* counters);

/* Batch over conditions with OpenMP (src/amici.cpp:235-237 restated).
 * Per-condition arrays are strided by their single-condition size. */
int amx_oracle_simulate_batch(const amx_oracle_model* m, const amx_oracle_opts* o,
                              int nb, const double* theta, const double* kfix,
                              const double* kpre, const double* x0, const double* sx0,
                              int nt, const double* ts,
                              const double* meas,
                              const double* sigma, double* x, double* y,
                              double* sx, double* llh, double* sllh,
                              long* counters, int* status, int nthreads);

/* Model pieces exposed for unit pinning. */
void amx_oracle_xdot(const amx_oracle_model* m, const double* p, const double* x,
                     double* xdot);
void amx_oracle_jac(const amx_oracle_model* m, const double* p, const double* x,
                    double* J /* nx*nx row-major */);
void amx_oracle_dxdotdp(const amx_oracle_model* m, const double* p,
                        const double* x, double* dfdp /* nx*np row-major */);
int amx_oracle_trace_forward(const amx_oracle_model* m, const amx_oracle_opts* o,
                             const double* theta, const double* kfix, int nt, const double* ts,
                             double* rec, int cap);
/* Called with the CVODES memory after every step of amx_oracle_trace_forward
 * when set (used by stald_probe.c to read the stability-limit data). */
extern void (*amx_oracle_step_hook)(void* cvode_mem);
double amx_oracle_Jy(int noise, double y, double sigma, double m);
double amx_oracle_dJydy(int noise, double y, double sigma, double m);

#ifdef __cplusplus
}
#endif
#endif
