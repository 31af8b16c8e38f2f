// Mass-action model evaluation, forward / forward-sensitivity / adjoint
// problems and the persistent batch kernels.  A generated model header defines
// a traits struct M (dimensions, team size and constexpr sparsity tables) and
// includes this file.
//
// Model semantics follow AMICI's SBML mass-action import, xdot = S w with
// w_r = p_r * prod(x_reactants) (python/sdist/amici/importers/sbml/__init__.py:620-627),
// the Jacobian / df/dp assembly of Model_ODE (src/model_ode.cpp:14-58, 141-170),
// sensitivity RHS (model_ode.cpp:426-450), adjoint RHS xBdot = -J^T xB and
// quadrature qBdot = -(df/dp)^T xB (model_ode.cpp:331-374), observables y = C x,
// Gaussian / log-normal Jy (python/sdist/amici/importers/utils.py:434-440),
// llh / sllh assembly (src/model.cpp:1194-1239, src/rdata.cpp:514-629) and the
// log10 chain rule (src/rdata.cpp:782-830).
#pragma once
#include "amx_bdf.h"

namespace amx {

// Launch shape of a model: teams per block chosen so that a block keeps at
// most 64 KiB of LDS and at most 256 threads.
template <class M>
struct Launch;

struct BatchArgs {
  int nb, nt;
  int ts_stride;           // 0: one time grid for all instances, else nt
  const double* theta;     // [nb][NP] parameters on their scale (pscale)
  const double* kfix;      // [nb][NK]
  const double* ts;        // [nb or 1][nt]
  const double* meas;      // [nb][nt][NY] (NaN = missing) or null
  const double* sigma;     // [nb][nt][NY]
  double* x;               // [nb][nt][NX] or null
  double* y;               // [nb][nt][NY] or null
  double* sx;              // [nb][nt][npl][NX] or null (FSA)
  double* llh;             // [nb]
  double* sllh;            // [nb][npl] (FSA / adjoint)
  const int* plist;        // [nplist] parameters of the sensitivities, or null (all, in order)
  int nplist;
  const int* pscale;       // [NP] ParameterScaling 0 none, 1 ln, 2 log10, or null (all log10)
  const double* x0;        // [nb][NX] initial states replacing the model's, or null
  const double* sx0;       // [nb][npl][NX] their sensitivities w.r.t. the scaled parameters, or null
  const double* t0;        // [nb] start times (ExpData t_start), or null (all 0)
  double* ts_rel;          // adjoint: [nb][nt] output times relative to the start time
  int* status;             // [nb]
  int* stats;              // [nb][AMX_NSTATS] or null
  int* diag;               // [nb][nt][AMX_NDIAG] per-output-time counters or null
  // adjoint forward trajectory (Hermite data at every accepted step; with
  // checkpoints, the steps of one checkpoint interval: max_traj = ck_every + 1)
  double* traj_t;          // [nb][max_traj]
  double* traj_y;          // [nb][max_traj][NX]
  double* traj_f;          // [nb][max_traj][NX] y' (Hermite interpolation) or null
  int* traj_q;             // [nb][max_traj] order used (polynomial interpolation) or null
  int* traj_n;             // [nb]
  double* dJydx;           // [nb][nt][NX]
  int max_traj;
  // adjoint checkpoints (ck_every > 0): a record of the integrator state every
  // ck_every forward steps, ck_pts forward points (steps + 1) at most
  int ck_every, ck_max, ck_pts;
  double* ck_rec;          // [nb][ck_max][CK_LEN]
  double* ck_spill;        // [nb][sizeof(BdfLds<NX, NP>) / 8] backward LDS during recomputation
  double* sens_ws;         // FSA workspace [nb][SENS_SLOTS][NP][NX]
  const double* kfix_pre;  // [nb][NK] pre-equilibration fixed parameters or null
  int newton_maxsteps;     // damped Newton iterations of the pre-equilibration
  double ss_rtol, ss_atol; // steady-state tolerances
  double* xss;             // adjoint: [nb][NX] pre-equilibrated steady state
  double* lu_ws;           // models with more than 64 states: [nb][NX][NX] Newton matrices
  int* counter;            // work-queue counter (zero before launch)
  int* first_fail;         // failfast: lowest failed instance (INT_MAX before launch)
  int failfast;            // skip instances after a failed one (run_simulations failfast)
  unsigned long long maxtime_ticks;  // per-instance wall-clock limit (0 = none)
  double* trace;           // optional step trace of one instance [cap][TRACE_REC]
  unsigned long long* prof;  // cycle profile [nb][2][PF_N] (diagnostic builds only)
  int trace_inst, trace_cap;
  SolverParams sp;         // forward
  SolverParams spB;        // backward (adjoint)

  // sensitivities computed (the length of the parameter list)
  __device__ __forceinline__ int npl(int np) const { return plist ? nplist : np; }
  __device__ __forceinline__ int pidx(int i) const { return plist ? plist[i] : i; }
  __device__ __forceinline__ int scale(int j) const { return pscale ? pscale[j] : 2; }
  __device__ __forceinline__ double tstart(int inst) const { return t0 ? t0[inst] : 0.0; }
};

// unscale_parameters (src/misc.cpp) for ParameterScaling none / ln / log10,
// and the chain-rule factor of ReturnData::apply_chain_rule_factor_to_
// simulation_results (src/rdata.cpp:782-796): p ln 10, p, 1
__device__ __forceinline__ double unscale_p(double v, int sc) {
  return sc == 2 ? exp10(v) : (sc == 1 ? exp(v) : v);
}
__device__ __forceinline__ double chain_coeff(double p, int sc) {
  return sc == 2 ? p * M_LN10 : (sc == 1 ? p : 1.0);
}

// The parameter list and scales are the caller's device arrays: an entry out
// of range fails the instance with AMICI_ILL_INPUT instead of being used
// (the reference raises on them in ConditionContext::apply_condition,
// src/edata.cpp:446-458, and Model::set_parameter_list).
template <class M, class TM>
__device__ __forceinline__ bool parameters_valid(const BatchArgs& a, const TM& tm) {
  bool bad = a.plist && (a.nplist < 0 || a.nplist > M::NP);
  if (a.plist && !bad)
    for (int i = tm.lane; i < a.nplist; i += M::T) bad |= a.plist[i] < 0 || a.plist[i] >= M::NP;
  if (a.pscale)
    for (int j = tm.lane; j < M::NP; j += M::T) bad |= a.pscale[j] < 0 || a.pscale[j] > 2;
  return !tm.any(bad);
}

// stats layout: forward nst, nfe, nsetups, nje, nni, netf, ncfn, last order,
//               backward nstB, nfeB, nsetupsB, njeB, nniB, netfB, ncfnB, nfQeB,
//               order reductions by stability-limit detection forward, backward
constexpr int AMX_NSTATS = 18;
// per-output-time diagnostics: forward numsteps, rhs evals, error-test
// failures, nonlinear-solver convergence failures, last order; backward
// numsteps, rhs evals, error-test failures, convergence failures
constexpr int AMX_NDIAG = 9;
constexpr int SENS_SLOTS = LMAX + 3;  // znS[0..5], acorS, ewtS, sllh partials

// Models with more than 64 states run one instance per workgroup (a
// multi-wavefront team): their Jacobian is kept as its sparse values Jv in
// LDS and the Newton matrix in global memory.
template <class M>
constexpr bool big_model() { return M::T > 64; }

template <class M>
struct Smem {
  static constexpr bool BIG = big_model<M>();
  static constexpr int JLD = M::NX + 1;
  static constexpr int NV = (M::NR > M::NY ? M::NR : M::NY) > M::NX
                                ? (M::NR > M::NY ? M::NR : M::NY) : M::NX;
  double x[M::NX + 1];  // x[NX] = 1: the factor of unimolecular table entries
  double v[M::NX];
  double pk[M::NR];   // rate constant of every reaction
  double w[NV];       // fluxes / monomials / observable scratch
  double u[NV];       // S^T v scratch
  double J[BIG ? 1 : M::NX * JLD];   // dense Jacobian (row-major)
  double Jv[BIG ? M::JNZ : 1];       // Jacobian values at (J_I[e], J_J[e])
};

template <class M>
struct Launch {
  static constexpr int BY_THREADS = 256 / M::T;
  static constexpr int BY_LDS =
      (int)(65536 / (sizeof(Smem<M>) + sizeof(BdfLds<M::NX, M::NP>)));
  static constexpr int TEAMS = big_model<M>() ? 1
                               : (BY_THREADS < BY_LDS ? BY_THREADS : BY_LDS) > 0
                                   ? (BY_THREADS < BY_LDS ? BY_THREADS : BY_LDS) : 1;
  static constexpr int THREADS = TEAMS * M::T;
  static_assert(M::T % 64 == 0 && THREADS <= 256, "whole wavefronts per team, <= 256 threads");
  // Register budget: models with at most 16 states are compiled for two
  // waves per SIMD (at most 256 registers per lane), so batches larger than
  // one instance per SIMD keep two resident per SIMD (4096 Robertson
  // conditions with forward sensitivities: 110 ms per batch at two, 171 ms at
  // one, 135 ms at four waves, where spills dominate); larger models keep the
  // full 512 registers of one wave per SIMD.  A module generated with
  // AMX_WAVES_PER_EU (a launch-shape variant, chosen by measurement) uses
  // that budget instead.
#ifdef AMX_WAVES_PER_EU
  static constexpr int WAVES_PER_EU = AMX_WAVES_PER_EU;
#else
  static constexpr int WAVES_PER_EU = M::NX <= 16 ? 2 : 1;
#endif
};

// ---------------------------------------------------------------------------
// model evaluation shared by all problems (team-collective)
template <class M>
struct MAEval {
  static constexpr int NX = M::NX, NR = M::NR, T = M::T;
  const Team<T>& tm;
  Smem<M>& s;

  __device__ __forceinline__ MAEval(const Team<T>& t, Smem<M>& sm) : tm(t), s(sm) {}

  __device__ __forceinline__ void put_x(double x) {
    team_sync();
    if (tm.lane < NX) s.x[tm.lane] = x;
    team_sync();
  }
  // monomials m_r = prod x_reactants into s.w (x already in s.x)
  __device__ __forceinline__ void monomials() {
    for (int r = tm.lane; r < NR; r += T) {
      const int a = M::RA[r], b = M::RB[r];
      double m = (a >= 0) ? s.x[a] : 1.0;
      if (b >= 0) m *= s.x[b];
      s.w[r] = m;
    }
    team_sync();
  }
  __device__ __forceinline__ double xdot(double x) {
    put_x(x);
    for (int r = tm.lane; r < NR; r += T) {
      const int a = M::RA[r], b = M::RB[r];
      double m = s.pk[r];
      if (a >= 0) m *= s.x[a];
      if (b >= 0) m *= s.x[b];
      s.w[r] = m;
    }
    team_sync();
    double f = 0.0;
#pragma unroll
    for (int q = 0; q < M::SPL_N; ++q) {
      const int k = q * T + tm.lane;
      f = fma(M::SPLC[k], s.w[M::SPL[k]], f);
    }
    return f;
  }
  // dense J into s.J (row-major, leading dim JLD), or for big models its
  // sparse values into s.Jv; x already in s.x
  __device__ __forceinline__ void jacobian() {
    constexpr int JLD = Smem<M>::JLD;
    if constexpr (!Smem<M>::BIG) {
      for (int e = tm.lane; e < NX * NX; e += T) s.J[(e / NX) * JLD + (e % NX)] = 0.0;
      team_sync();
    }
    for (int e = tm.lane; e < M::JNZ; e += T) {
      double v = 0.0;
      for (int t = M::JT_PTR[e]; t < M::JT_PTR[e + 1]; ++t) {
        double d = M::JT_C[t] * s.pk[M::JT_R[t]];
        const int o = M::JT_O[t];
        if (o >= 0) d *= s.x[o];
        v += d;
      }
      if constexpr (Smem<M>::BIG) s.Jv[e] = v;
      else s.J[M::J_I[e] * JLD + M::J_J[e]] = v;
    }
    team_sync();
  }
  // d I + g J (TRANS: d I + g J^T) of a big model into the column-major
  // global matrix A (row i, column j at A[j * NX + i]): the Newton matrix of
  // the integrator (d = 1) or the Jacobian itself (d = 0, g = 1)
  template <bool TRANS>
  __device__ __forceinline__ void newton_matrix(double g, double* A, double d = 1.0) {
    for (int e = tm.lane; e < NX * NX; e += T) {
      const int j = e / NX, i = e - j * NX;
      A[e] = (i == j) ? d : 0.0;
    }
    __syncthreads();  // global entries written by one lane, updated by another
    for (int e = tm.lane; e < M::JNZ; e += T) {
      const int r = TRANS ? M::J_J[e] : M::J_I[e];
      const int c = TRANS ? M::J_I[e] : M::J_J[e];
      A[(size_t)c * NX + r] += g * s.Jv[e];
    }
    __syncthreads();
  }
  // u_r = sum_i S_ir v_i into s.u; v in s.v.  The per-lane sums of this and
  // the other model evaluations run over the lane-major padded tables (entry q
  // of all lanes in one coalesced row, fixed trip counts, fully unrolled), so
  // every table load of a sum is issued at once.
  __device__ __forceinline__ void stoich_t() {
    constexpr int ER = (NR + T - 1) / T;
#pragma unroll
    for (int e = 0; e < ER; ++e) {
      const int r = tm.lane + e * T;
      double u = 0.0;
#pragma unroll
      for (int q = 0; q < M::RSL_N; ++q) {
        const int k = (e * M::RSL_N + q) * T + tm.lane;
        u = fma(M::RSLC[k], s.v[M::RSL[k]], u);
      }
      if (r < NR) s.u[r] = u;
    }
    team_sync();
  }
};

// ---------------------------------------------------------------------------
// forward problem, optionally with NS staggered forward sensitivities
template <class M, int NS_, bool CK = false>
struct FwdProblem {
  static constexpr int NX = M::NX, T = M::T, NQ = 0, NS = NS_, NP = M::NP;
  // checkpoints: the integrator writes a record when ck_due >= 0 and takes a
  // fresh Newton matrix after it
  static constexpr bool FORCE_SETUP = CK, CK_SAVE = CK;
  static constexpr int JLD = Smem<M>::JLD;
  const Team<T>& tm;
  Smem<M>& s;
  MAEval<M> ev;
  double* ws;  // FSA workspace of this instance: [SENS_SLOTS][NS][NX]
  // sensitivities computed: ns <= NS, sensitivity is with respect to
  // parameter pidx(is) (the parameter list; all parameters in order without one)
  int ns = NS;
  const int* plist = nullptr;
  __device__ __forceinline__ int pidx(int is) const { return plist ? plist[is] : is; }
  // adjoint forward pass: Hermite data of every accepted step
  double* tt;
  double* ty;
  double* tf;
  int* tq;
  int ntraj, max_traj;
  // checkpoints (CK): every ck_K points the record ck_due = point / ck_K is
  // due, and only the current interval's points are kept (slot point % ck_K);
  // a recomputation (ck_base >= 0) stores the points ck_base .. ck_base + ck_K
  // in slots 0 .. ck_K and then stops
  int ck_K = 0, ck_base = -1, ck_due = -1, jout = 0;
  double* ck_rec = nullptr;

  __device__ __forceinline__ FwdProblem(const Team<T>& t, Smem<M>& sm, double* w)
      : tm(t), s(sm), ev(t, sm), ws(w), tt(nullptr), ty(nullptr), tf(nullptr), tq(nullptr),
        ntraj(0), max_traj(0) {}

  // the forward point after every accepted step (and the initial point):
  // t, y and, for Hermite interpolation, y' = zn[1] / h
  // (CVAhermiteStorePnt, cvodea.c:2839-2895), for polynomial interpolation the
  // order used (CVApolynomialStorePnt, cvodea.c:3236-3262; 0 at the initial
  // point)
  __device__ __forceinline__ bool on_step(double tn, double z0, double z1, double hs, int qu) {
    if (tt == nullptr) return true;
    int slot = ntraj;
    if constexpr (CK) {
      if (ck_base >= 0) {
        slot = ntraj - ck_base;
      } else {
        if (ntraj >= max_traj) return false;
        slot = ntraj % ck_K;
        if (ntraj > 0 && slot == 0) ck_due = ntraj / ck_K;  // CVodeF, cvodea.c:577-597
      }
    } else {
      if (ntraj >= max_traj) return false;
    }
    if (tm.lane == 0) {
      tt[slot] = tn;
      if (tq) tq[slot] = qu;
    }
    if (tm.lane < NX) {
      ty[(size_t)slot * NX + tm.lane] = z0;
      if (tf) tf[(size_t)slot * NX + tm.lane] = z1 / hs;
    }
    ntraj++;
    if constexpr (CK) {
      if (ck_base >= 0 && ntraj - ck_base > ck_K) return false;  // interval complete
    }
    return true;
  }

  __device__ __forceinline__ double rhs(double, double y) { return ev.xdot(y); }
  __device__ __forceinline__ void quad_rhs(double, double, double*) {}
  __device__ __forceinline__ void jac(double, double y) {
    ev.put_x(y);
    ev.jacobian();
  }
  __device__ __forceinline__ void build_M(double gamma, double* a) {
    const int i = tm.lane;
#pragma unroll
    for (int j = 0; j < NX; ++j)
      a[j] = (i < NX) ? ((i == j ? 1.0 : 0.0) - gamma * s.J[i * JLD + j]) : 0.0;
  }
  double* mat = nullptr;  // big models: this instance's Newton matrix (global)
  __device__ __forceinline__ void build_M_global(double gamma) {
    ev.template newton_matrix<false>(-gamma, mat);
  }

  // ---------- forward sensitivities (global-memory Nordsieck arrays) ----------
  __device__ __forceinline__ double& S(int slot, int is) const {
    const int i = tm.lane < NX ? tm.lane : NX - 1;
    return ws[((size_t)slot * NS + is) * NX + i];
  }
  static constexpr int ACOR = LMAX, EWT = LMAX + 1, ACC = LMAX + 2;
  double ssp_rtol, ssp_atol;

  // Stages y (for J*v through the flux structure) and the monomials (df/dp).
  __device__ __forceinline__ void sens_prepare(double y) {
    ev.put_x(y);
    ev.monomials();
  }
  // sensitivity RHS for column is (parameter j = pidx(is)) at the staged y:
  //   (J v)_i + df_i/dp_j = sum_r S_ir (dw_r/dx . v + [KI(r) == j] m_r)
  __device__ __forceinline__ double sens_f(double yS, int is) {
    const int pj = pidx(is);
    team_sync();
    if (tm.lane < NX) s.v[tm.lane] = yS;
    team_sync();
    for (int r = tm.lane; r < M::NR; r += T) {
      const int a = M::RA[r], b = M::RB[r];
      double d = 0.0;
      if (a >= 0) d = s.v[a] * ((b >= 0) ? s.x[b] : 1.0);
      if (b >= 0) d = fma(s.v[b], s.x[a], d);
      double u = s.pk[r] * d;
      if (M::KI[r] == pj) u += s.w[r];
      s.u[r] = u;
    }
    team_sync();
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < M::SPL_N; ++q) {
      const int k = q * T + tm.lane;
      acc = fma(M::SPLC[k], s.u[M::SPL[k]], acc);
    }
    return acc;
  }
  __device__ __forceinline__ double sens_norm_slot(int slot) const {
    double m = 0.0;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double e = (tm.lane < NX) ? S(slot, is) * S(EWT, is) : 0.0;
      m = fmax(m, sqrt(tm.sum(e * e) / NX));
    }
    return m;
  }
  __device__ __forceinline__ void sens_init() {
    if (tm.lane < NX)
      for (int sl = 0; sl < SENS_SLOTS; ++sl)
        for (int is = 0; is < NS && is < ns; ++is) S(sl, is) = 0.0;  // sx0 = dx0/dp = 0
  }
  __device__ __forceinline__ void sens_set_ewt() {
    if (tm.lane < NX)
      for (int is = 0; is < NS && is < ns; ++is) S(EWT, is) = 1.0 / (ssp_rtol * fabs(S(0, is)) + ssp_atol);
  }
  __device__ __forceinline__ void sens_predict(int q) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      double z[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? S(j, is) : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] += z[j];
#pragma unroll
      for (int j = 0; j < QMAX; ++j)
        if (j < q) S(j, is) = z[j];
    }
  }
  __device__ __forceinline__ void sens_restore(int q) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      double z[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? S(j, is) : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] -= z[j];
#pragma unroll
      for (int j = 0; j < QMAX; ++j)
        if (j < q) S(j, is) = z[j];
    }
  }
  __device__ __forceinline__ void sens_rescale(int q, double eta) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      double c = eta;
#pragma unroll
      for (int j = 1; j <= QMAX; ++j) {
        if (j <= q) S(j, is) *= c;
        c *= eta;
      }
    }
  }
  __device__ __forceinline__ void sens_increase(int q, double A1, const double (&ll)[LMAX]) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double zL = A1 * S(QMAX, is);
#pragma unroll
      for (int j = 2; j <= QMAX; ++j)
        if (j <= q) S(j, is) = fma(ll[j], zL, S(j, is));
      S(q + 1, is) = zL;
    }
  }
  __device__ __forceinline__ void sens_decrease(int q, const double (&ll)[LMAX]) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double zq = S(q, is);
#pragma unroll
      for (int j = 2; j < QMAX; ++j)
        if (j < q) S(j, is) = fma(-ll[j], zq, S(j, is));
    }
  }
  __device__ __forceinline__ void sens_complete(int q, const double (&l)[LMAX]) {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double a = S(ACOR, is);
#pragma unroll
      for (int j = 0; j <= QMAX; ++j)
        if (j <= q) S(j, is) = fma(l[j], a, S(j, is));
    }
  }
  __device__ __forceinline__ void sens_save_acor() {
    if (tm.lane >= NX) return;
    for (int is = 0; is < NS && is < ns; ++is) S(QMAX, is) = S(ACOR, is);
  }
  __device__ __forceinline__ double sens_norm_zn(int j) { return sens_norm_slot(j); }
  __device__ __forceinline__ double sens_norm_qp1(double cquot) {
    double m = 0.0;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double e = (tm.lane < NX) ? (S(ACOR, is) - cquot * S(QMAX, is)) * S(EWT, is) : 0.0;
      m = fmax(m, sqrt(tm.sum(e * e) / NX));
    }
    return m;
  }
  __device__ __forceinline__ double sens_hub_inv() {
    double m = 0.0;
    if (tm.lane < NX)
      for (int is = 0; is < NS && is < ns; ++is)
        m = fmax(m, fabs(S(1, is)) / (0.1 * fabs(S(0, is)) + 1.0 / S(EWT, is)));
    return tm.max(m);
  }
  // znS[1] = fS(t, y, znS[0]) (unscaled)
  __device__ __forceinline__ void sens_init_deriv(double, double y, double) {
    sens_prepare(y);
    for (int is = 0; is < NS && is < ns; ++is) {
      const double v = sens_f(tm.lane < NX ? S(0, is) : 0.0, is);
      if (tm.lane < NX) S(1, is) = v;
    }
  }
  __device__ __forceinline__ void sens_scale_deriv(double h) {
    if (tm.lane < NX)
      for (int is = 0; is < NS && is < ns; ++is) S(1, is) *= h;
  }
  __device__ __forceinline__ void sens_restart_q1(double t, double y, double f, double h) {
    sens_init_deriv(t, y, f);
    sens_scale_deriv(h);
  }
  double ydd_hg;
  __device__ __forceinline__ void sens_ydd_prepare(double hg) { ydd_hg = hg; }
  __device__ __forceinline__ double sens_ydd_norm(double, double yy, double, double hg) {
    sens_prepare(yy);
    double m = 0.0;
    for (int is = 0; is < NS && is < ns; ++is) {
      const double ys = (tm.lane < NX) ? S(0, is) + hg * S(1, is) : 0.0;
      const double fs = sens_f(ys, is);
      const double e = (tm.lane < NX) ? (fs - S(1, is)) / hg * S(EWT, is) : 0.0;
      m = fmax(m, sqrt(tm.sum(e * e) / NX));
    }
    return m;
  }

  // cvStgrNls: Newton on all sensitivity corrections with the state's
  // Newton matrix (no setup unless the iteration fails with a stale J).
  template <class B>
  __device__ __forceinline__ int sens_nls(B& b) {
    sens_prepare(b.y);
    bool callSetup = false, jbad = false;
    for (;;) {
      if (tm.lane < NX)
        for (int is = 0; is < NS && is < ns; ++is) S(ACOR, is) = 0.0;
      if (callSetup) {
        if (jbad) b.convfail = FAIL_BAD_J;
        const int r = b.lsetup(b.convfail);
        sens_prepare(b.y);
        if (r != NLS_OK) return NLS_RECOVER;
      }
      int m = 0;
      int ret = NLS_OK;
      for (;;) {
        b.nniS++;
        double Del = 0.0;
        for (int is = 0; is < NS && is < ns; ++is) {
          const double acS = (tm.lane < NX) ? S(ACOR, is) : 0.0;
          const double yS = (tm.lane < NX) ? S(0, is) + acS : 0.0;
          const double fS = sens_f(yS, is);
          const double res = (tm.lane < NX) ? b.rl1 * S(1, is) + acS - b.gamma * fS : 0.0;
          const double d = b.lsolve(-res);
          if (tm.lane < NX) S(ACOR, is) = acS + d;
          const double e = (tm.lane < NX) ? d * S(EWT, is) : 0.0;
          Del = fmax(Del, sqrt(tm.sum(e * e) / NX));
        }
        if (m > 0) b.crateS = fmax(CRDOWN * b.crateS, Del / b.delp);
        const double dcon = Del * fmin(1.0, b.crateS) / b.S.tq[4];
        if (dcon <= 1.0) {
          b.acnrmS = (m == 0) ? Del : sens_norm_slot(ACOR);
          b.jcur = false;
          return NLS_OK;
        }
        if (m >= 1 && Del > RDIV * b.delp) { ret = NLS_RECOVER; break; }
        b.delp = Del;
        m++;
        if (m >= NLS_MAXCOR) { ret = NLS_RECOVER; break; }
      }
      if (ret == NLS_RECOVER && !b.jcur) {
        callSetup = true;
        jbad = true;
        continue;
      }
      return ret;
    }
  }
};

// ---------------------------------------------------------------------------
// Forward solution x(t) during the adjoint sweep, from the forward steps
// stored by the forward pass.  POLY (AMICI's default, CV_POLYNOMIAL): the
// Newton form of the polynomial through order + 1 stored points, order being
// the order used for the step that ends the interval, with divided
// differences scaled by the interval length (CVApolynomialGetY,
// cvodea.c:3270-3380); the interval search restates CVAfindIndex
// (cvodea.c): t in (t[i-1], t[i]].  Otherwise cubic Hermite interpolation
// between the interval's end points from the stored y and y'
// (CVAhermiteGetY, cvodea.c:2906).
template <class M, bool POLY>
struct TrajInterp;

template <class M>
struct TrajInterp<M, true> {
  static constexpr int NX = M::NX;
  static constexpr int QM = LMAX;  // at most order 5: six points
  const double* tt;  // forward step times (increasing)
  const double* ty;  // [n][NX]
  const int* tq;     // [n] order of the step ending at each stored time
  int n, ilast, idx, order;
  double dt, T[QM], Y[QM];  // points and divided differences (this lane's component)

  __device__ __forceinline__ void init(const double* tt_, const double* ty_, const int* tq_,
                                       const double*, int n_) {
    tt = tt_;
    ty = ty_;
    tq = tq_;
    n = n_;
    ilast = n - 1;
    idx = -1;
    order = 0;
    dt = 1.0;
#pragma unroll
    for (int j = 0; j < QM; ++j) T[j] = Y[j] = 0.0;
  }
  __device__ __forceinline__ double eval(double t, int lane) {
    const int i = lane < NX ? lane : 0;
    if (n < 2) return ty[i];
    // CVAfindIndex: interval index with t in (tt[k-1], tt[k]]
    int k = ilast;
    bool newpoint = idx < 0;
    if (t < tt[ilast - 1]) {
      newpoint = true;
      for (int g = 0; g < n && k > 0 && t <= tt[k - 1]; ++g) --k;
      ilast = (k == 0) ? 1 : k;
    } else if (t > tt[ilast]) {
      newpoint = true;
      for (int g = 0; g < n && k < n - 1 && t > tt[k]; ++g) ++k;
      ilast = k;
    }
    if (k == 0) return ty[i];  // left limit
    if (newpoint || k != idx) {
      idx = k;
      dt = fabs(tt[k] - tt[k - 1]);
      int base = k;
      order = tq[base];
      if (k < order) base += order - k;
#pragma unroll
      for (int j = 0; j < QM; ++j) {
        if (j <= order) {
          T[j] = tt[base - j];
          Y[j] = ty[(size_t)(base - j) * NX + i];
        }
      }
#pragma unroll
      for (int a = 1; a < QM; ++a) {
#pragma unroll
        for (int j = QM - 1; j >= a; --j) {
          if (a <= order && j <= order) {
            const double f = dt / (T[j] - T[j - a]);
            Y[j] = f * (Y[j] - Y[j - 1]);  // N_VLinearSum(f, Y[j], -f, Y[j-1]): f (x - y)
          }
        }
      }
    }
    // nested products c_0 = 1, c_{j+1} = c_j (t - T_j) / dt; x = sum_j c_j Y_j
    double c = 1.0, x = Y[0];
#pragma unroll
    for (int j = 0; j < QM - 1; ++j) {
      if (j < order) {
        c = c * (t - T[j]) / dt;
        x = fma(c, Y[j + 1], x);
      }
    }
    return x;
  }
};

template <class M>
struct TrajInterp<M, false> {
  static constexpr int NX = M::NX;
  const double* tt;  // forward step times (increasing)
  const double* ty;  // [n][NX]
  const double* tf;  // [n][NX] y'
  int n, cur;        // interval [cur-1, cur]
  double ta, tb, ya, yb, fa, fb;

  __device__ __forceinline__ void init(const double* tt_, const double* ty_, const int*,
                                       const double* tf_, int n_) {
    tt = tt_;
    ty = ty_;
    tf = tf_;
    n = n_;
    cur = -1;
    ta = tb = ya = yb = fa = fb = 0.0;
  }
  __device__ __forceinline__ void fetch(int k, int i, double& t_, double& y_, double& f_) const {
    t_ = tt[k];
    y_ = ty[(size_t)k * NX + i];
    f_ = tf[(size_t)k * NX + i];
  }
  // Selects the interval [tt[k-1], tt[k]]; a move to the previous interval
  // keeps the shared end point and loads only the new start point.
  __device__ __forceinline__ void set_interval(int k, int i) {
    if (cur == k + 1) {
      tb = ta;
      yb = ya;
      fb = fa;
    } else {
      fetch(k, i, tb, yb, fb);
    }
    fetch(k - 1, i, ta, ya, fa);
    cur = k;
  }
  __device__ __forceinline__ double eval(double t, int lane) {
    const int i = lane < NX ? lane : 0;
    if (n < 2) return ty[i];
    if (cur <= 0) set_interval(n - 1, i);
    // the backward sweep moves down through the stored steps; after a failed
    // step it may move up again (bounded walks: at most n intervals)
    for (int g = 0; g < n && cur > 1 && t < ta; ++g) set_interval(cur - 1, i);
    for (int g = 0; g < n && cur < n - 1 && t > tb; ++g) set_interval(cur + 1, i);
    const double hk = tb - ta;
    const double sx = (t - ta) / hk;
    const double om = 1.0 - sx;
    const double h00 = (1.0 + 2.0 * sx) * om * om;
    const double h10 = sx * om * om;
    const double h01 = sx * sx * (3.0 - 2.0 * sx);
    const double h11 = sx * sx * (sx - 1.0);
    return h00 * ya + h10 * hk * fa + h01 * yb + h11 * hk * fb;
  }
};

// adjoint problem: xB' = -J(x(t))^T xB, qB' = -(df/dp)^T xB, x(t) from the
// stored forward steps (TrajInterp)
template <class M, bool POLY, bool CK = false>
struct BwdProblem {
  static constexpr int NX = M::NX, T = M::T, NQ = M::NP, NS = 0, NP = M::NP;
  // checkpoints: a fresh Newton matrix after every recomputation, which
  // reuses this problem's Jacobian storage
  static constexpr bool FORCE_SETUP = CK;
  static constexpr int JLD = Smem<M>::JLD;
  static constexpr int EPQ = (NQ + T - 1) / T;
  const Team<T>& tm;
  Smem<M>& s;
  MAEval<M> ev;
  TrajInterp<M, POLY> ip;
  double t_cached, x_cached;
  bool have_cached;
  // quadratures integrated: nq <= NQ; quadrature slot j belongs to parameter
  // pidx(j) (the parameter list; all parameters in order without one)
  int nq = NQ;
  const int* plist = nullptr;
  __device__ __forceinline__ int pidx(int j) const { return plist ? plist[j] : j; }

  __device__ __forceinline__ bool on_step(double, double, double, double, int) { return true; }

  __device__ __forceinline__ BwdProblem(const Team<T>& t, Smem<M>& sm, const double* tt_,
                                        const double* ty_, const int* tq_, const double* tf_,
                                        int n_)
      : tm(t), s(sm), ev(t, sm) {
    ip.init(tt_, ty_, tq_, tf_, n_);
    t_cached = 0.0;
    x_cached = 0.0;
    have_cached = false;
  }

  __device__ __forceinline__ double xf(double t) {
    if (have_cached && t == t_cached) return x_cached;
    x_cached = ip.eval(t, tm.lane);
    t_cached = t;
    have_cached = true;
    return x_cached;
  }

  // puts x(t) and the adjoint state into LDS and computes u = S^T xB
  __device__ __forceinline__ void stage(double t, double yB) {
    const double x = xf(t);
    team_sync();
    if (tm.lane < NX) {
      s.x[tm.lane] = x;
      s.v[tm.lane] = yB;
    }
    team_sync();
    ev.stoich_t();
  }

  __device__ __forceinline__ double rhs(double t, double yB) {
    stage(t, yB);
    double f = 0.0;
#pragma unroll
    for (int q = 0; q < M::DXL_N; ++q) {
      const int k = q * T + tm.lane;
      const unsigned e = M::DXL[k];
      const int r = e & 0xFFFF, o = e >> 16;
      f = fma(-M::DXLC[k] * s.pk[r] * s.x[o], s.u[r], f);
    }
    return f;
  }

  __device__ __forceinline__ void quad_rhs(double t, double yB, double* fq) {
    stage(t, yB);
    ev.monomials();
#pragma unroll
    for (int e = 0; e < EPQ; ++e) {
      // the entries of parameter pj in the parameter-lane tables: lane pj % T
      // of pass pj / T (the lane itself when there is no parameter list)
      const int j = tm.lane + e * T;
      const int pj = pidx(j < nq ? j : 0);
      const int base = (pj / T) * M::PRL_N * T + pj % T;
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < M::PRL_N; ++q) {
        const int k = base + q * T;
        const int r = M::PRL[k];
        acc = fma(-M::PRLC[k] * s.w[r], s.u[r], acc);
      }
      fq[e] = (j < nq) ? acc : 0.0;
    }
  }

  __device__ __forceinline__ void jac(double t, double) {
    const double x = xf(t);
    ev.put_x(x);
    ev.jacobian();
  }
  // Newton matrix of the backward problem: I - gamma * (-J^T)
  __device__ __forceinline__ void build_M(double gamma, double* a) {
    const int i = tm.lane;
#pragma unroll
    for (int j = 0; j < NX; ++j)
      a[j] = (i < NX) ? ((i == j ? 1.0 : 0.0) + gamma * s.J[j * JLD + i]) : 0.0;
  }
  double* mat = nullptr;  // big models: this instance's Newton matrix (global)
  __device__ __forceinline__ void build_M_global(double gamma) {
    ev.template newton_matrix<true>(gamma, mat);
  }
};

// ---------------------------------------------------------------------------
// observables and objective at one output time; returns this team's Jy sum
// contribution (already reduced) and leaves dJy/dx of this lane in dJdx.
struct ObjOut {
  double nllh;  // team sum of Jy at this output time
  double dJdx;  // dJy/dx of this lane's state
};
template <class M>
__device__ __forceinline__ ObjOut objective_at(const Team<M::T>& tm, Smem<M>& s, double x,
                                               const double* meas, const double* sigma,
                                               double* yout) {
  constexpr int NY = M::NY, NX = M::NX, T = M::T;
  team_sync();
  if (tm.lane < NX) s.x[tm.lane] = x;
  team_sync();
  double nllh = 0.0;
  for (int k = tm.lane; k < NY; k += T) {
    double yk = 0.0;
    for (int q = M::OB_PTR[k]; q < M::OB_PTR[k + 1]; ++q) yk = fma(M::OB_C[q], s.x[M::OB_S[q]], yk);
    if (yout) yout[k] = yk;
    double g = 0.0;
    const double mv = meas ? meas[k] : NAN;
    if (!isnan(mv)) {
      const double sg = sigma[k];
      // Jy and dJy/dy of the noise model (importers/utils.py:434-461): r is
      // the difference of y and m on the model's scale (linear, ln or log10)
      // divided by sigma, and dr/dy = 1 / dd with dd = sigma, sigma y or
      // sigma y ln 10
      constexpr int NZ = M::NOISE;
      constexpr bool LAP = NZ >= 3;
      constexpr int SC = NZ % 3;  // 0 linear, 1 ln, 2 log10 (NOISE ids 0..5)
      double r, dd, jc;
      if (SC == 0) {
        r = (yk - mv) / sg;
        dd = sg;
        jc = LAP ? log(2.0 * sg) : 0.5 * log(2.0 * M_PI * sg * sg);
      } else if (SC == 1) {
        r = (log(yk) - log(mv)) / sg;
        dd = sg * yk;
        jc = LAP ? log(2.0 * sg * mv) : 0.5 * log(2.0 * M_PI * sg * sg * mv * mv);
      } else {
        r = (log10(yk) - log10(mv)) / sg;
        dd = sg * yk * M_LN10;
        jc = LAP ? log(2.0 * sg * mv * M_LN10)
                 : 0.5 * log(2.0 * M_PI * sg * sg * mv * mv * M_LN10 * M_LN10);
      }
      if (LAP) {
        nllh += jc + fabs(r);
        g = (r > 0.0 ? 1.0 : (r < 0.0 ? -1.0 : 0.0)) / dd;
      } else {
        nllh += jc + 0.5 * r * r;
        g = r / dd;
      }
    }
    s.w[k] = g;
  }
  team_sync();
  double d = 0.0;
  if (tm.lane < NX)
    for (int q = M::OT_PTR[tm.lane]; q < M::OT_PTR[tm.lane + 1]; ++q)
      d = fma(M::OT_C[q], s.w[M::OT_Y[q]], d);
  team_sync();
  return ObjOut{tm.sum(nllh), d};
}

__device__ __forceinline__ bool next_t_too_close(double cur, double next) {
  const double tdiff = next - cur;
  if (tdiff == 0.0) return true;
  const double tround = UROUND * fmax(fabs(cur), fabs(next));
  return fabs(tdiff) < 2.0 * tround;
}

template <class B>
__device__ __forceinline__ void store_stats(int* st, const B& b, int off) {
  st[off + 0] = b.nst;
  st[off + 1] = b.nfe;
  st[off + 2] = b.nsetups;
  st[off + 3] = b.nje;
  st[off + 4] = b.nni;
  st[off + 5] = b.netf;
  st[off + 6] = b.ncfn;
  st[off + 7] = (off == 0) ? b.q : b.nfQe;
  st[16 + (off == 0 ? 0 : 1)] = b.nor;
}

// ---------------------------------------------------------------------------
// pre-equilibration to steady state (SteadyStateProblem,
// src/forwardproblem.cpp:755-1223; NewtonsMethod::run,
// src/steadystateproblem.cpp:74-161)

// WRMS of xdot with the steady-state tolerances (WRMSComputer,
// src/steadystateproblem.cpp:42-57); convergence when < 1.
template <class M>
__device__ __forceinline__ double ss_wrms(const Team<M::T>& tm, double f, double x,
                                          const BatchArgs& a) {
  const double e = (tm.lane < M::NX) ? f / (a.ss_rtol * fabs(x) + a.ss_atol) : 0.0;
  return sqrt(tm.sum(e * e) / M::NX);
}

// Factorises J(x) (rows of J^T when TRANS) into the LU of `lds`.
template <class M, bool TRANS, class L>
__device__ __forceinline__ bool factor_jacobian(const Team<M::T>& tm, Smem<M>& s, MAEval<M>& ev,
                                                double x, LURows<M::NX, M::T>& lu, L& lds) {
  constexpr int NX = M::NX, JLD = Smem<M>::JLD;
  ev.put_x(x);
  ev.jacobian();
  const bool act = tm.lane < NX;
  const int cl = act ? tm.lane : 0;
  double row[NX];
#pragma unroll
  for (int j = 0; j < NX; ++j)
    row[j] = act ? (TRANS ? s.J[j * JLD + cl] : s.J[cl * JLD + j]) : 0.0;
  return lu.factor(tm, row, lds);
}

// LU of the Jacobian J(x) (or J^T) for the steady-state Newton steps and the
// implicit-function sensitivities: row-per-lane in LDS for one-wavefront
// teams, the blocked global-memory LU (matrix in `mat`) for workgroup teams.
template <class M, bool BIG = big_model<M>()>
struct JacLU {
  LURows<M::NX, M::T> lu;
  double* mat = nullptr;
  template <bool TRANS, class L>
  __device__ __forceinline__ bool factor(const Team<M::T>& tm, Smem<M>& s, MAEval<M>& ev, double x,
                                         L& lds) {
    return factor_jacobian<M, TRANS>(tm, s, ev, x, lu, lds);
  }
  template <class L>
  __device__ __forceinline__ double solve(double b, const Team<M::T>& tm, L& lds) const {
    return lu.solve(b, tm, lds);
  }
};
template <class M>
struct JacLU<M, true> {
  LUBig<M::NX, M::T> lu;
  double* mat = nullptr;
  template <bool TRANS, class L>
  __device__ __forceinline__ bool factor(const Team<M::T>& tm, Smem<M>&, MAEval<M>& ev, double x,
                                         L& lds) {
    ev.put_x(x);
    ev.jacobian();
    ev.template newton_matrix<TRANS>(1.0, mat, 0.0);
    return lu.factor(tm, mat, lds);
  }
  template <class L>
  __device__ __forceinline__ double solve(double b, const Team<M::T>& tm, L& lds) const {
    return lu.solve(b, tm, mat, lds);
  }
};

// Damped Newton on xdot(x) = 0 from x: the direction solves J delta = -xdot
// (NewtonSolver::get_step), the damping factor doubles after a step that
// lowers the residual and is quartered otherwise, failing below 1e-8
// (NewtonsMethod::update_damping_factor).  If Newton does not converge
// within newton_maxsteps, the initial state is integrated over horizons
// 1, 10, 100, ... until the criterion holds.  On success x is the steady
// state; otherwise an AMICI status is returned.
template <class M>
__device__ __forceinline__ int steady_state(const BatchArgs& a, const Team<M::T>& tm, Smem<M>& s,
                                            BdfLds<M::NX, 0>& lds, double& x, double* mat) {
  constexpr int NX = M::NX;
  MAEval<M> ev(tm, s);
  const bool act = tm.lane < NX;
  const double x0 = x;
  double f = ev.xdot(x);
  double best = ss_wrms<M>(tm, f, x, a);
  if (best < 1.0) return AMX_SUCCESS;
  int status = AMX_NO_STEADY_STATE;
  {
    JacLU<M> jl;
    jl.mat = mat;
    double xold = x, delta = 0.0, gam = 1.0;
    bool ok = true;
    for (int it = 0; it < a.newton_maxsteps; ++it) {
      if (ok) {
        xold = x;
        if (!jl.template factor<false>(tm, s, ev, x, lds)) {
          status = AMX_SINGULAR_JACOBIAN;
          break;
        }
        delta = jl.solve(act ? -f : 0.0, tm, lds);
        if (!act) delta = 0.0;
      }
      x = xold + gam * delta;
      f = ev.xdot(x);
      const double w = ss_wrms<M>(tm, f, x, a);
      ok = w < best;
      if (ok) {
        best = w;
        if (w < 1.0) return AMX_SUCCESS;
      }
      gam = ok ? fmin(1.0, 2.0 * gam) : 0.25 * gam;
      if (gam < 1e-8) {
        status = AMX_DAMPING_FACTOR_ERROR;
        break;
      }
    }
  }
  // integration from the initial state
  FwdProblem<M, 0> pp(tm, s, nullptr);
  pp.mat = mat;
  BDF<FwdProblem<M, 0>> bdf(pp, lds, a.sp);
  if (a.maxtime_ticks) bdf.sp.deadline = wall_clock64() + a.maxtime_ticks;
  bdf.init(0.0, x0, nullptr);
  double tout = 1.0;
  for (int k = 0; k < 16; ++k, tout *= 10.0) {
    const int st = bdf.advance(tout, tout, a.sp.maxsteps - bdf.nst);
    if (st != AMX_SUCCESS) return status == AMX_NO_STEADY_STATE ? st : status;
    const double xv = bdf.yout;
    const double fv = ev.xdot(xv);
    if (ss_wrms<M>(tm, fv, xv, a) < 1.0) {
      x = xv;
      return AMX_SUCCESS;
    }
  }
  return AMX_NO_STEADY_STATE;
}

// ---------------------------------------------------------------------------
// forward simulation of one instance.  MODE 0: states, 1: FSA, 2: adjoint
// forward pass (stores the trajectory and dJy/dx at every output), 3: adjoint
// forward pass with checkpoints (a record of the integrator state every
// ck_every steps, the last interval's points).
template <class M, int MODE>
__device__ __forceinline__ void simulate_forward(int inst, const BatchArgs& a, const Team<M::T>& tm,
                                                 Smem<M>& s, BdfLds<M::NX, 0>& lds) {
  constexpr int NX = M::NX, NP = M::NP, NK = M::NK, NY = M::NY, T = M::T;
  constexpr int NS = (MODE == 1) ? NP : 0;
  constexpr bool ADJ = MODE >= 2, CKM = MODE == 3;
  const size_t I = (size_t)inst;
  for (int r = tm.lane; r < M::NR; r += T)
    s.pk[r] = unscale_p(a.theta[I * NP + M::KI[r]], a.scale(M::KI[r]));
  if (tm.lane == 0) s.x[M::NX] = 1.0;
  team_sync();
  // initial state: ExpData::x0 (Model::set_initial_state) or the model's
  // (constants and fixed parameters)
  double x0 = 0.0;
  if (tm.lane < NX) {
    const int kx = M::X0K[tm.lane];
    x0 = a.x0 ? a.x0[I * NX + tm.lane] : ((kx >= 0) ? a.kfix[I * NK + kx] : M::X0C[tm.lane]);
  }
  // pre-equilibration: steady state at the pre-equilibration inputs, then the
  // fixed-parameter-dependent initial states are reset from kfix
  // (reinitialize_fixed_parameter_initial_states)
  int pre_status = parameters_valid<M>(a, tm) ? AMX_SUCCESS : AMX_ILL_INPUT;
  double xss = 0.0;
  constexpr bool BIG = big_model<M>();
  double* const mat = BIG ? a.lu_ws + I * (size_t)NX * NX : nullptr;
  if (a.kfix_pre && pre_status == AMX_SUCCESS) {
    // from ExpData::x0 if given, else from the model's initial state at the
    // pre-equilibration inputs (Model::initialize_states)
    double xp = 0.0;
    if (tm.lane < NX) {
      const int kx = M::X0K[tm.lane];
      xp = (a.x0 || kx < 0) ? x0 : a.kfix_pre[I * NK + kx];
    }
    pre_status = steady_state<M>(a, tm, s, lds, xp, mat);
    xss = xp;
    if (tm.lane < NX) {
      const int kx = M::X0K[tm.lane];
      x0 = (kx >= 0) ? a.kfix[I * NK + kx] : xp;
    }
    if constexpr (ADJ) {
      if (tm.lane < NX) a.xss[I * NX + tm.lane] = xp;
    }
  }
  double* ws = (MODE == 1) ? a.sens_ws + I * (size_t)SENS_SLOTS * NP * NX : nullptr;
  using PB = FwdProblem<M, NS, CKM>;
  PB pb0(tm, s, ws);
  pb0.ns = (MODE == 1) ? a.npl(NP) : 0;
  pb0.plist = a.plist;
  const int npl = a.npl(NP);
  pb0.ssp_rtol = a.sp.rtolS;
  pb0.ssp_atol = a.sp.atolS;
  if constexpr (ADJ) {
    pb0.tt = a.traj_t + I * (size_t)a.max_traj;
    pb0.ty = a.traj_y + I * (size_t)a.max_traj * NX;
    pb0.tf = a.traj_f ? a.traj_f + I * (size_t)a.max_traj * NX : nullptr;
    pb0.tq = a.traj_q ? a.traj_q + I * (size_t)a.max_traj : nullptr;
    pb0.max_traj = a.max_traj;
  }
  if constexpr (CKM) {
    pb0.max_traj = a.ck_pts;
    pb0.ck_K = a.ck_every;
    pb0.ck_rec = a.ck_rec + I * (size_t)a.ck_max * BDF<PB>::CK_LEN;
  }
  pb0.mat = mat;
  PB& pb = pb0;
  BDF<PB> bdf(pb, lds, a.sp);
  if (a.maxtime_ticks) bdf.sp.deadline = wall_clock64() + a.maxtime_ticks;
  // The models are autonomous, so a start time t0 (ExpData t_start) only
  // shifts the output grid: the integration runs in the time relative to t0,
  // from 0, and an output time at or before t0 gives the initial state.  The
  // adjoint passes read the relative output times the forward pass stores.
  const double t0 = a.tstart(inst);
  bdf.init(0.0, x0, nullptr);
  if constexpr (CKM) bdf.save_state(pb.ck_rec, 0);  // checkpoint 0: the initial state
  if (a.trace && inst == a.trace_inst) {
    bdf.trace = a.trace;
    bdf.trace_cap = a.trace_cap / 2;  // second half: the backward problem's records
  }
  if constexpr (MODE == 1) {
    pb.sens_init();
    // ExpData::sx0 (Model::set_initial_state_sensitivities, src/model.cpp:986-1023):
    // given w.r.t. the scaled parameters, integrated divided by the chain-rule factor
    if (a.sx0 && tm.lane < NX)
      for (int is = 0; is < NS && is < pb.ns; ++is) {
        const int pj = pb.pidx(is);
        const int sc = a.scale(pj);
        const double cf = chain_coeff(unscale_p(a.theta[I * NP + pj], sc), sc);
        pb.S(0, is) = a.sx0[(I * npl + is) * NX + tm.lane] / cf;
      }
  }
  const double* ts = a.ts + (size_t)a.ts_stride * I;

  int status = pre_status;
  if constexpr (MODE == 1) {
    // steady-state sensitivities by the implicit function theorem:
    // sx0 = -J(x_ss)^{-1} df/dp for the species not reset from kfix
    if (a.kfix_pre && status == AMX_SUCCESS) {
      JacLU<M> jl;
      jl.mat = mat;
      if (!jl.template factor<false>(tm, s, pb.ev, xss, lds)) {
        status = AMX_SINGULAR_JACOBIAN;
      } else {
        pb.sens_prepare(xss);
        for (int is = 0; is < pb.ns; ++is) {
          const double fp = pb.sens_f(0.0, is);
          const double v = jl.solve(tm.lane < NX ? -fp : 0.0, tm, lds);
          if (tm.lane < NX) pb.S(0, is) = (M::X0K[tm.lane] >= 0) ? 0.0 : v;
        }
      }
    }
  }
  double llh = 0.0;
  int it = 0;
  for (; it < a.nt && status == AMX_SUCCESS; ++it) {
    const double tout = fmax(ts[it] - t0, 0.0);
    if constexpr (ADJ) {
      if (tm.lane == 0) a.ts_rel[I * a.nt + it] = tout;
    }
    if (tout > 0.0 && bdf.tn < tout && !next_t_too_close(bdf.tn, tout)) {
      // the step budget counts over the whole forward run; one used up
      // exactly at an output time is an exception in Solver::apply_max_num_steps
      // (src/solver.cpp:107-119), reported as AMICI_ERROR (src/amici.cpp:156-157)
      if (bdf.nst >= bdf.sp.maxsteps) {
        status = AMX_ERROR;
        break;
      }
      if constexpr (CKM) pb.jout = it;
      status = bdf.advance(tout, tout, bdf.sp.maxsteps - bdf.nst);
      if (status != AMX_SUCCESS) break;
    }
    const double x = bdf.yout;
    if (a.diag && tm.lane == 0) {
      // Solver::store_diagnosis (src/solver.cpp:349-375): cumulative counters,
      // zeros before the solver was first called
      int* d = a.diag + (I * a.nt + it) * AMX_NDIAG;
      const bool called = bdf.nst > 0;
      d[0] = called ? bdf.nst : 0;
      d[1] = called ? bdf.nfe : 0;
      d[2] = called ? bdf.netf : 0;
      d[3] = called ? bdf.ncfn : 0;
      d[4] = called ? bdf.q : 0;
      for (int k = 5; k < AMX_NDIAG; ++k) d[k] = 0;  // backward slots (adjoint sweep)
    }
    if (a.x && tm.lane < NX) a.x[(I * a.nt + it) * NX + tm.lane] = x;
    const size_t mo = (I * a.nt + it) * NY;
    const ObjOut ob = objective_at<M>(tm, s, x, a.meas ? a.meas + mo : nullptr, a.sigma + mo,
                                      a.y ? a.y + mo : nullptr);
    llh -= ob.nllh;
    const double dJdx = ob.dJdx;
    if constexpr (ADJ) {
      if (tm.lane < NX) a.dJydx[(I * a.nt + it) * NX + tm.lane] = dJdx;
    }
    if constexpr (MODE == 1) {
      if (tm.lane < NX)
        for (int is = 0; is < pb.ns; ++is) {
          const double sxv = pb.S(0, is);
          pb.S(PB::ACC, is) -= dJdx * sxv;
          if (a.sx) a.sx[((I * a.nt + it) * npl + is) * NX + tm.lane] = sxv;
        }
    }
  }
  if (status != AMX_SUCCESS) {
    llh = NAN;
    if (a.diag && tm.lane == 0)
      for (int j = it; j < a.nt; ++j)
        for (int k = 0; k < AMX_NDIAG; ++k) a.diag[(I * a.nt + j) * AMX_NDIAG + k] = 0;
    for (int j = it; j < a.nt; ++j) {
      if (a.x && tm.lane < NX) a.x[(I * a.nt + j) * NX + tm.lane] = NAN;
      if (a.y)
        for (int k = tm.lane; k < NY; k += T) a.y[(I * a.nt + j) * NY + k] = NAN;
    }
  }
#ifdef AMX_PROFILE
  if (a.prof && tm.lane == 0)
    for (int k = 0; k < PF_N; ++k) a.prof[I * 2 * PF_N + k] = bdf.prof[k];
#endif
  if (tm.lane == 0) {
    a.llh[inst] = llh;
    a.status[inst] = status;
    if (a.stats) store_stats(a.stats + I * AMX_NSTATS, bdf, 0);
    if constexpr (ADJ) a.traj_n[inst] = pb.ntraj;
  }
  if constexpr (MODE == 1) {
    // sllh and sx: chain rule of each parameter's scale on the non-zero
    // entries (rdata.cpp:782-866)
    for (int is = 0; is < pb.ns; ++is) {
      const double part = tm.sum(tm.lane < NX ? pb.S(PB::ACC, is) : 0.0);
      const int pj = pb.pidx(is);
      const int sc = a.scale(pj);
      const double cf = chain_coeff(unscale_p(a.theta[I * NP + pj], sc), sc);
      const double v = (status == AMX_SUCCESS) ? (part != 0.0 ? part * cf : 0.0) : NAN;
      if (tm.lane == 0) a.sllh[I * npl + is] = v;
      if (a.sx && tm.lane < NX)
        for (int j = 0; j < a.nt; ++j) {
          double& e = a.sx[((I * a.nt + j) * npl + is) * NX + tm.lane];
          e = (j < it) ? (e != 0.0 ? e * cf : 0.0) : NAN;
        }
    }
  }
}

// Recomputation of checkpoint interval c of one instance (CVAdataStore,
// cvodea.c:2328-2375): the forward problem restarts from record c and stores
// the points c * K .. (c + 1) * K in the trajectory buffer.  Record 0 is the
// initial state, so interval 0 repeats the start of the forward pass; a later
// record restarts inside the forward pass's call towards output jout with a
// fresh Newton matrix, as the forward pass did after writing it, so the
// recomputed steps are the forward pass's.  The problem's LDS block is the
// backward problem's, which the caller saved.
// With a step trace (trace != nullptr) the recomputed steps are appended to
// it from record trace_n on.
template <class M>
__device__ __forceinline__ int recompute_interval(const BatchArgs& a, const Team<M::T>& tm,
                                                           Smem<M>& s, BdfLds<M::NX, 0>& lds,
                                                           int inst, int c,
                                                           unsigned long long deadline,
                                                           double* trace, int trace_cap,
                                                           int& trace_n) {
  constexpr int NX = M::NX;
  using PR = FwdProblem<M, 0, true>;
  const size_t I = (size_t)inst;
  PR pr(tm, s, nullptr);
  pr.ns = 0;
  pr.tt = a.traj_t + I * (size_t)a.max_traj;
  pr.ty = a.traj_y + I * (size_t)a.max_traj * NX;
  pr.tf = a.traj_f ? a.traj_f + I * (size_t)a.max_traj * NX : nullptr;
  pr.tq = a.traj_q ? a.traj_q + I * (size_t)a.max_traj : nullptr;
  pr.max_traj = a.max_traj;
  pr.ck_K = a.ck_every;
  pr.ck_base = c * a.ck_every;
  pr.ntraj = pr.ck_base;
  pr.mat = big_model<M>() ? a.lu_ws + I * (size_t)NX * NX : nullptr;
  BDF<PR> bdf(pr, lds, a.sp);
  bdf.sp.deadline = deadline;
  bdf.trace = trace;
  bdf.trace_cap = trace_cap;
  bdf.trace_n = trace_n;
  const int jout = bdf.load_state(a.ck_rec + (I * (size_t)a.ck_max + c) * BDF<PR>::CK_LEN);
  if (c > 0) {
    bdf.force_setup = true;
    pr.on_step(bdf.tn, bdf.Z(0), bdf.Z(1), bdf.hscale, bdf.q);
  }
  const double* ts = a.ts_rel + I * a.nt;  // as the forward pass saw them
  int r = AMX_ERR_FAILURE;  // the forward pass went on past this interval
  for (int j = jout; j < a.nt; ++j) {
    const double tout = ts[j];
    if (tout > 0.0 && bdf.tn < tout && !next_t_too_close(bdf.tn, tout)) {
      const int rr = bdf.advance(tout, tout, 0x3fffffff);
      if (rr == AMX_TRAJ_FULL) {  // point (c + 1) * K stored
        r = AMX_SUCCESS;
        break;
      }
      if (rr != AMX_SUCCESS) {
        r = rr;
        break;
      }
    }
  }
  trace_n = bdf.trace_n;
  return r;
}

// backward (adjoint) sweep of one instance; EventHandlingBwdSimulator::run
// (src/backwardproblem.cpp:267-332) with a fresh order-1 restart after every
// data-point jump (CVodeSolver::reinit_b, src/solver_cvodes.cpp:594-649).
// With checkpoints (CK) the sweep runs interval by interval as CVodeB does
// (cvodea.c:1490-1600): each call integrates with the stop time at the
// current interval's start, and the interval before it is recomputed when the
// sweep reaches that start.  lds_f is the same LDS block seen as the forward
// problem's; the backward problem's contents are kept in ck_spill meanwhile.
template <class M, bool POLY, bool CK = false>
__device__ __forceinline__ void simulate_backward(int inst, const BatchArgs& a, const Team<M::T>& tm,
                                                  Smem<M>& s, BdfLds<M::NX, M::NP>& lds,
                                                  BdfLds<M::NX, 0>* lds_f = nullptr) {
  constexpr int NX = M::NX, NP = M::NP, T = M::T;
  using PB = BwdProblem<M, POLY, CK>;
  constexpr int EPQ = PB::EPQ;
  const size_t I = (size_t)inst;
  int* const diag = a.diag ? a.diag + I * a.nt * AMX_NDIAG : nullptr;
  if (diag && tm.lane == 0)
    for (int j = 0; j < a.nt; ++j)
      for (int k = 5; k < AMX_NDIAG; ++k) diag[(size_t)j * AMX_NDIAG + k] = 0;
  const int npl = a.npl(NP);
  const int fst = a.status[inst];
  if (fst != AMX_SUCCESS) {
    for (int j = tm.lane; j < npl; j += T) a.sllh[I * npl + j] = NAN;
    return;
  }
  for (int r = tm.lane; r < M::NR; r += T)
    s.pk[r] = unscale_p(a.theta[I * NP + M::KI[r]], a.scale(M::KI[r]));
  if (tm.lane == 0) s.x[M::NX] = 1.0;
  team_sync();
  const int n = a.traj_n[inst];
  const double* tt = a.traj_t + I * (size_t)a.max_traj;
  const double* ty = a.traj_y + I * (size_t)a.max_traj * NX;
  const int* tqo = POLY ? a.traj_q + I * (size_t)a.max_traj : nullptr;
  const double* tfd = POLY ? nullptr : a.traj_f + I * (size_t)a.max_traj * NX;
  // checkpoints: interval c spans the points c * K .. (c + 1) * K; the
  // forward pass left the last one, of n_last points, in the buffer
  const int K = CK ? a.ck_every : 0;
  int c = 0, n_last = n;
  if constexpr (CK) {
    c = (n > 0) ? (n - 1) / K : 0;
    n_last = n - c * K;
  }
  PB pb0(tm, s, tt, ty, tqo, tfd, n_last);
  pb0.nq = npl;
  pb0.plist = a.plist;
  constexpr bool BIG = big_model<M>();
  if constexpr (BIG) pb0.mat = a.lu_ws + I * (size_t)NX * NX;
  BDF<PB> bdf(pb0, lds, a.spB);  // holds pb0 and the LDS block by reference
  if (a.maxtime_ticks) bdf.sp.deadline = wall_clock64() + a.maxtime_ticks;
  // step trace: the backward steps in the second half of the buffer; with
  // checkpoints the second half's first half, and the recomputed forward
  // steps in its second half
  double* rtrace = nullptr;
  int rtrace_cap = 0, rtrace_n = 0;
  if (a.trace && inst == a.trace_inst) {
    bdf.trace = a.trace + (size_t)(a.trace_cap / 2) * TRACE_REC;
    bdf.trace_cap = a.trace_cap - a.trace_cap / 2;
    if constexpr (CK) {
      rtrace_cap = bdf.trace_cap - bdf.trace_cap / 2;
      bdf.trace_cap /= 2;
      rtrace = bdf.trace + (size_t)bdf.trace_cap * TRACE_REC;
    }
  }
  const double* ts = a.ts_rel + I * a.nt;  // relative to the start time (forward pass)
  const double* dJ = a.dJydx + I * a.nt * NX;
  const double t0 = 0.0;
  int it = a.nt - 1;
  double tcur = (n > 0) ? tt[n_last - 1] : t0;
  double xB = 0.0;
  double qB[EPQ];
#pragma unroll
  for (int e = 0; e < EPQ; ++e) qB[e] = 0.0;
  // per-output-time backward diagnostics (Solver::store_diagnosis_b,
  // src/solver.cpp:377-398, called from src/backwardproblem.cpp:222-290):
  // entry k at the k-th data point from the last one, zeros until the
  // backward solver has been called
  int ndiag = 0;
  if (it >= 0 && ts[it] == tcur) {
    if (tm.lane < NX) xB += dJ[(size_t)it * NX + tm.lane];
    --it;
    ++ndiag;
  }
  int status = AMX_SUCCESS;
  // EventHandlingBwdSimulator::run (src/backwardproblem.cpp:267-332) with
  // CVodeB (cvodea.c:1365-1630): the backward integration stops only at the
  // checkpoint start t0 (CVodeSetStopTime(ck_t0)), so steps may pass a data
  // point, whose adjoint state is then the Nordsieck interpolant; after the
  // jump the problem restarts at order 1 (reset_state) and, once it has been
  // integrated before, takes its first step through reInit_post_process with
  // the stop time at the next data point.
  const int mxB = a.spB.maxsteps;
  // start time of checkpoint interval k (record k's time; t0 for k = 0)
  const double* ckr = CK ? a.ck_rec + I * (size_t)a.ck_max * BDF<PB>::CK_LEN : nullptr;
  auto ck_t = [&](int k) __attribute__((always_inline)) -> double {
    return k > 0 ? ckr[(size_t)k * BDF<PB>::CK_LEN + LMAX * NX + 7 + 6 + 6 + 15] : t0;
  };
  // moves to the interval before the current one while the backward problem
  // is at or before the current one's start, recomputing it
  auto enter = [&]() __attribute__((always_inline)) -> int {
    if constexpr (CK) {
      for (int g = 0; g < n && c > 0 && bdf.tn <= ck_t(c); ++g) {
        --c;
        double* spill = a.ck_spill + I * (size_t)(sizeof(BdfLds<NX, NP>) / sizeof(double));
        double* l64 = reinterpret_cast<double*>(&lds);
        constexpr int NW = (int)(sizeof(BdfLds<NX, NP>) / sizeof(double));
        team_sync();
        for (int k = tm.lane; k < NW; k += T) spill[k] = l64[k];
        team_sync();
        const int r = recompute_interval<M>(a, tm, s, *lds_f, inst, c, bdf.sp.deadline, rtrace,
                                            rtrace_cap, rtrace_n);
        team_sync();
        for (int k = tm.lane; k < NW; k += T) l64[k] = spill[k];
        team_sync();
        if (r != AMX_SUCCESS) return r;
        pb0.ip.init(tt, ty, tqo, tfd, K + 1);
        pb0.have_cached = false;
        bdf.force_setup = true;
      }
    }
    return AMX_SUCCESS;
  };
  // CVode towards tnext, across checkpoint intervals (CK)
  auto advance_to = [&](double tnext) __attribute__((always_inline)) -> int {
    for (int g = 0; g <= n; ++g) {
      int r = enter();
      if (r != AMX_SUCCESS) return r;
      const double tst = ck_t(c);
      r = bdf.advance(tnext, tst, mxB);
      if (r != AMX_SUCCESS) return r;
      if (!bdf.at_tstop || tst <= tnext) return AMX_SUCCESS;
    }
    return AMX_ERR_FAILURE;
  };
  auto reinit_to = [&](double tnext) __attribute__((always_inline)) -> int {
    const int r = enter();
    if (r != AMX_SUCCESS) return r;
    const double tst = ck_t(c) > tnext ? ck_t(c) : tnext;
    return bdf.reinit_step(tnext, tst, mxB);
  };
  bool called = false, force_pp = false;
  bdf.init(tcur, xB, qB);
  while (it >= 0 && status == AMX_SUCCESS) {
    const double tnext = ts[it];
    if (tnext < tcur) {
      if (force_pp) {
        if constexpr (CK) status = reinit_to(tnext);
        else status = bdf.reinit_step(tnext, mxB);
        force_pp = false;
        if (status != AMX_SUCCESS) break;
      }
      if constexpr (CK) status = advance_to(tnext);
      else status = bdf.advance(tnext, t0, mxB);
      if (status != AMX_SUCCESS) break;
      called = true;
      tcur = tnext;
      xB = bdf.yout;
#pragma unroll
      for (int e = 0; e < EPQ; ++e) qB[e] = bdf.yQout[e];
    }
    if (diag && tm.lane == 0 && ndiag < a.nt) {
      int* d = diag + (size_t)ndiag * AMX_NDIAG;
      d[5] = called ? bdf.nst : 0;
      d[6] = called ? bdf.nfe : 0;
      d[7] = called ? bdf.netf : 0;
      d[8] = called ? bdf.ncfn : 0;
    }
    ++ndiag;
    if (tm.lane < NX) xB += dJ[(size_t)it * NX + tm.lane];
    --it;
    if (called) {
      force_pp = true;
      bdf.reset_state(tcur, xB, qB);
    } else {
      bdf.init(tcur, xB, qB);
    }
  }
  if (status == AMX_SUCCESS && tcur > t0) {
    if (force_pp) {
      if constexpr (CK) status = reinit_to(t0);
      else status = bdf.reinit_step(t0, mxB);
    }
    if (status == AMX_SUCCESS) {
      if constexpr (CK) status = advance_to(t0);
      else status = bdf.advance(t0, t0, mxB);
    }
    if (status == AMX_SUCCESS) {
      xB = bdf.yout;
#pragma unroll
      for (int e = 0; e < EPQ; ++e) qB[e] = bdf.yQout[e];
    }
  }
  // pre-equilibrated instance: sllh also gets -xB(t0) . sx0 with the
  // implicit-function sensitivities sx0 = -J(x_ss)^{-1} df/dp of the species
  // not reset from kfix, i.e. + v . df/dp with J(x_ss)^T v = xB(t0) on those
  // species (SteadyStateBackwardProblem::compute_quadrature_by_lin_solve,
  // src/backwardproblem.cpp:428-460)
  double gss[EPQ];
#pragma unroll
  for (int e = 0; e < EPQ; ++e) gss[e] = 0.0;
  if (a.kfix_pre && status == AMX_SUCCESS) {
    const double xs = a.xss[I * NX + (tm.lane < NX ? tm.lane : 0)];
    const double xbm = (tm.lane < NX && M::X0K[tm.lane] < 0) ? xB : 0.0;
    JacLU<M> jl;
    jl.mat = pb0.mat;
    if (!jl.template factor<true>(tm, s, pb0.ev, xs, lds)) {
      status = AMX_SINGULAR_JACOBIAN;
    } else {
      double v = jl.solve(xbm, tm, lds);
      if (tm.lane >= NX) v = 0.0;
      team_sync();
      if (tm.lane < NX) s.v[tm.lane] = v;
      team_sync();
      pb0.ev.monomials();
      pb0.ev.stoich_t();
#pragma unroll
      for (int e = 0; e < EPQ; ++e) {
        const int j = tm.lane + e * T;
        double acc = 0.0;
        if (j < npl) {
          const int pj = pb0.pidx(j);
          for (int q = M::PR_PTR[pj]; q < M::PR_PTR[pj + 1]; ++q) {
            const int r = M::PR_R[q];
            acc = fma(s.w[r], s.u[r], acc);
          }
        }
        gss[e] = acc;
      }
    }
  }
  // caller-given initial state sensitivities (ExpData::sx0, w.r.t. the scaled
  // parameters): -xB(t0) . sx0_j / (chain-rule factor of parameter j)
  double gsx[EPQ];
#pragma unroll
  for (int e = 0; e < EPQ; ++e) gsx[e] = 0.0;
  if (a.sx0) {
    team_sync();
    if (tm.lane < NX) s.v[tm.lane] = xB;
    team_sync();
#pragma unroll
    for (int e = 0; e < EPQ; ++e) {
      const int j = tm.lane + e * T;
      if (j < npl) {
        const double* sj = a.sx0 + (I * npl + j) * NX;
        double d = 0.0;
        for (int i = 0; i < NX; ++i) d = fma(s.v[i], sj[i], d);
        gsx[e] = d;
      }
    }
    team_sync();
  }
  // sllh = -(xB(t0).sx0 + qB); chain rule of each parameter's scale on the
  // non-zero entries
#pragma unroll
  for (int e = 0; e < EPQ; ++e) {
    const int j = tm.lane + e * T;
    if (j < npl) {
      const int pj = pb0.pidx(j);
      const int sc = a.scale(pj);
      const double cf = chain_coeff(unscale_p(a.theta[I * NP + pj], sc), sc);
      const double g = gss[e] - qB[e] - gsx[e] / cf;
      a.sllh[I * npl + j] = (status == AMX_SUCCESS) ? (g != 0.0 ? g * cf : 0.0) : NAN;
    }
  }
#ifdef AMX_PROFILE
  if (a.prof && tm.lane == 0)
    for (int k = 0; k < PF_N; ++k) a.prof[(I * 2 + 1) * PF_N + k] = bdf.prof[k];
#endif
  if (tm.lane == 0) {
    if (status != AMX_SUCCESS) a.status[inst] = status;
    if (a.stats) store_stats(a.stats + I * AMX_NSTATS, bdf, 8);
  }
}

// ---------------------------------------------------------------------------
// persistent kernels: every team pulls instance indices from a device counter
template <class M>
__device__ __forceinline__ int next_instance(const Team<M::T>& tm, int* counter) {
  int v = 0;
  if (tm.lane == 0) v = atomicAdd(counter, 1);
  return tm.bcast_i(v, 0);
}

// failfast (run_simulations' skipThrough, src/amici.cpp:232-260): once an
// instance has failed, instances with a larger index that have not started
// are not run (status AMICI_NOT_RUN, llh and sllh NaN); the caller reports
// every instance after the first failed one as not run, so the results do
// not depend on which of them had already started
template <class M, class TM>
__device__ __forceinline__ bool skip_failfast(const BatchArgs& a, const TM& tm, int inst) {
  if (!a.failfast) return false;
  int f = 0;
  if (tm.lane == 0) f = __hip_atomic_load(a.first_fail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  f = tm.bcast_i(f, 0);
  if (inst <= f) return false;
  const size_t I = (size_t)inst;
  const int npl = a.npl(M::NP);
  if (tm.lane == 0) {
    a.status[inst] = AMX_NOT_RUN;
    a.llh[inst] = NAN;
    if (a.stats)
      for (int k = 0; k < AMX_NSTATS; ++k) a.stats[I * AMX_NSTATS + k] = 0;
  }
  if (a.sllh)
    for (int j = tm.lane; j < npl; j += M::T) a.sllh[I * npl + j] = NAN;
  return true;
}
template <class TM>
__device__ __forceinline__ void note_failure(const BatchArgs& a, const TM& tm, int inst) {
  if (a.failfast && tm.lane == 0 && a.status[inst] < 0) atomicMin(a.first_fail, inst);
}

template <class M, int MODE>
__device__ __forceinline__ void forward_kernel_body(BatchArgs a) {
  constexpr int T = M::T;
  __shared__ Smem<M> smem[Launch<M>::TEAMS];
  __shared__ BdfLds<M::NX, 0> blds[Launch<M>::TEAMS];
  Team<T> tm;
  Smem<M>& s = smem[threadIdx.x / T];
  BdfLds<M::NX, 0>& lds = blds[threadIdx.x / T];
  for (int guard = 0; guard <= a.nb; ++guard) {
    const int inst = next_instance<M>(tm, a.counter);
    if (inst >= a.nb) break;
    if (skip_failfast<M>(a, tm, inst)) continue;
    simulate_forward<M, MODE>(inst, a, tm, s, lds);
    note_failure(a, tm, inst);
  }
}

template <class M, bool POLY>
__device__ __forceinline__ void backward_kernel_body(BatchArgs a) {
  constexpr int T = M::T;
  __shared__ Smem<M> smem[Launch<M>::TEAMS];
  __shared__ BdfLds<M::NX, M::NP> blds[Launch<M>::TEAMS];
  Team<T> tm;
  Smem<M>& s = smem[threadIdx.x / T];
  BdfLds<M::NX, M::NP>& lds = blds[threadIdx.x / T];
  for (int guard = 0; guard <= a.nb; ++guard) {
    const int inst = next_instance<M>(tm, a.counter);
    if (inst >= a.nb) break;
    if (skip_failfast<M>(a, tm, inst)) continue;
    simulate_backward<M, POLY>(inst, a, tm, s, lds);
    note_failure(a, tm, inst);
  }
}

// backward sweep with checkpoints: the LDS block of a team serves the
// backward problem and, during a recomputation, the forward problem
template <class M, bool POLY>
__device__ __forceinline__ void backward_ck_kernel_body(BatchArgs a) {
  constexpr int T = M::T;
  union Blk {
    BdfLds<M::NX, M::NP> b;
    BdfLds<M::NX, 0> f;
  };
  static_assert(sizeof(BdfLds<M::NX, 0>) <= sizeof(BdfLds<M::NX, M::NP>), "LDS block");
  __shared__ Smem<M> smem[Launch<M>::TEAMS];
  __shared__ Blk blds[Launch<M>::TEAMS];
  Team<T> tm;
  Smem<M>& s = smem[threadIdx.x / T];
  Blk& blk = blds[threadIdx.x / T];
  for (int guard = 0; guard <= a.nb; ++guard) {
    const int inst = next_instance<M>(tm, a.counter);
    if (inst >= a.nb) break;
    if (skip_failfast<M>(a, tm, inst)) continue;
    simulate_backward<M, POLY, true>(inst, a, tm, s, blk.b, &blk.f);
    note_failure(a, tm, inst);
  }
}

}  // namespace amx
