// Variable-order BDF integrator for one ODE instance per team of lanes
// (gfx950, FP64).  Each team lane i < NX owns state component i.  The step,
// order and Newton heuristics follow SUNDIALS CVODES 7.1.1 as AMICI drives it
// (BDF, Newton, dense direct solver, tstop at every output time):
//   cvStep                       ThirdParty/sundials/src/cvodes/cvodes.c:5800-6094
//   cvAdjustParams/Order/BDF     :6112-6423      cvRescale       :6425
//   cvPredict                    :6507           cvSetBDF/TqBDF  :6788-6880
//   cvNls + Newton               :6894, sunnonlinsol_newton.c, cvodes_nls.c:244-353
//   cvLsSetup / cvLsSolve        cvodes_ls.c:1611-1853 (J reuse, 2/(1+gamrat) scaling)
//   cvHandleNFlag / cvRestore    cvodes.c:7309-7400
//   cvDoErrorTest                :7461           cvCompleteStep  :7634
//   cvPrepareNextStep/SetEta     :7747-7830      cvComputeEtaqm1/qp1, cvChooseEta :7836-8010
//   cvHin / cvUpperBoundH0 / cvYddNorm :5436-5798 CVode tstop handling :3100-3620
//   stability-limit detection    cvBDFStab :8158-8219, cvSLdet :8256 (amx_stald.h)
//   constants                    cvodes_impl.h:57-204, cvodes_ls_impl.h:40-42
//   quadratures (cvQuadNls)      cvodes.c:7089;  staggered FSA (cvStgrNls) :7167,
//                                cvodes_nls_stg.c (residual / lsolve / conv test)
// Dense LU with partial pivoting replaces SUNDlsMat_denseGETRF/GETRS
// (sundials_dense.c:82-148): row i of the Newton matrix belongs to lane i, is
// factorised in that lane's registers and kept in LDS between setup and
// solves; the pivot order is a permutation instead of row swaps.
// Storage: the Nordsieck history, the LU factors and the method coefficient
// arrays of an instance live in LDS (BdfLds); scalars and per-lane work values
// live in registers.  Every function is inlined into the batch kernels.
// Not modelled: hmin/hmax limits (AMICI sets neither).
#pragma once
#include <type_traits>

#include "amx_stald.h"
#include "amx_team.h"

namespace amx {

constexpr int QMAX = 5;
constexpr int LMAX = QMAX + 1;
constexpr double UROUND = 2.220446049250313e-16;

// CVODES tuning constants.
constexpr double ETA_MAX_FX = 1.5, ETA_MAX_FS = 10000.0, ETA_MAX_ES = 10.0,
                 ETA_MAX_GS = 10.0, ETA_MIN = 0.1, ETA_MAX_EF = 0.2,
                 ETA_MIN_EF = 0.1, ETA_CF = 0.25;
constexpr int SMALL_NST = 10, SMALL_NEF = 2, MXNEF1 = 3, LONG_WAIT = 10,
              MXNCF = 10, MXNEF = 7, NLS_MAXCOR = 3, MSBP = 20, MSBJ = 51;
// Upper bound on the attempts of one step before the failure counters stop
// it: the state / quadrature counters (ncf < MXNCF, nef, nefQ < MXNEF) are
// reset each time the sensitivity stage is entered, and that stage allows
// ncfS < MXNCF, nefS < MXNEF, so at most (MXNCF + MXNEF - 1) *
// (MXNCF + 2 MXNEF - 2) attempts occur; the guard sits above that bound and
// only ends a step that the counters would not end.
constexpr int MAX_ATTEMPTS = (MXNCF + MXNEF) * (MXNCF + 2 * MXNEF);
constexpr double ADDON = 1e-6, BIAS1 = 6.0, BIAS2 = 6.0, BIAS3 = 10.0,
                 DGMAX_LSETUP = 0.3, DGMAX_JBAD = 0.2, CRDOWN = 0.3,
                 RDIV = 2.0, CORTES = 0.1, FUZZ_FACTOR = 100.0,
                 HLB_FACTOR = 100.0, HUB_FACTOR = 0.1, H_BIAS = 0.5;

// AMICI status codes (include/amici/defines.h:66-99) plus one of this library.
enum : int {
  AMX_SUCCESS = 0,
  AMX_TOO_MUCH_WORK = -1,
  AMX_TOO_MUCH_ACC = -2,
  AMX_ERR_FAILURE = -3,
  AMX_CONV_FAILURE = -4,
  AMX_LSETUP_FAIL = -6,
  AMX_RHSFUNC_FAIL = -8,
  AMX_ILL_INPUT = -22,
  AMX_TOO_CLOSE = -27,
  AMX_ERROR = -99,
  AMX_MAX_TIME_EXCEEDED = -1000,
  AMX_NO_STEADY_STATE = -81,
  AMX_DAMPING_FACTOR_ERROR = -86,
  AMX_SINGULAR_JACOBIAN = -9987,
  AMX_NOT_RUN = -1001,
  AMX_TRAJ_FULL = -1010,  // forward-trajectory buffer of an instance exhausted
};

// internal flags
enum : int { FIRST_CALL = 0, PREV_CONV_FAIL = 1, PREV_ERR_FAIL = 2 };
enum : int { DO_ERROR_TEST = 10, PREDICT_AGAIN = 11, TRY_AGAIN = 12 };
enum : int { NLS_OK = 0, NLS_RECOVER = 1 };
enum : int { NO_FAILURES = 0, FAIL_BAD_J = 1, FAIL_OTHER = 2 };

__device__ __forceinline__ double powr(double b, double e) { return pow(b, e); }

// a[i] for a runtime index i
template <int N>
__device__ __forceinline__ double sel(const double (&a)[N], int i) {
  double r = a[0];
#pragma unroll
  for (int k = 1; k < N; ++k)
    if (i == k) r = a[k];
  return r;
}

// LDS block of one integrator instance.  Models with more than 64 states
// keep their LU factors in global memory (LUBig) instead.
template <int NX, int NQ>
struct BdfLds {
  static constexpr int NQP = NQ > 0 ? NQ : 1;
  static constexpr int NLU = NX <= 64 ? NX : 1;
  double zn[LMAX][NX];    // Nordsieck history of the states (column i: lane i)
  double znQ[LMAX][NQP];  // Nordsieck history of the quadratures
  double lu[NLU][NLU];    // LU factors: (row i, column j) at lu[j][i]
  double tau[LMAX + 1];   // previous step sizes
  double l[LMAX];         // method coefficients
  double tq[6];           // error-test constants
  unsigned int permw[(NLU + 3) / 4];  // lane of the k-th pivot row: byte k % 4 of word k / 4
  double wsol[NLU];       // solve: y_k, then x_k, of pivot step k
  double ssdat[5][3];     // stability-limit detection: squared scaled derivative norms
                          // of orders q-1, q, q+1 of the last five steps, newest first
};

// Row-per-lane dense LU of an NX x NX matrix held by the first NX team lanes.
// The factors are kept in LDS (BdfLds::lu, permw); the per-lane pivot
// bookkeeping lives in registers.  Neither the factorisation nor the solve
// compares the lane index with a loop constant (such per-step lane masks are
// loop invariant, and the compiler keeps them in scalar registers for the
// whole kernel): the pivot order is packed four steps to a 32-bit word, the
// pivot search is a maximum reduction and a wave ballot, and the
// substitutions move values between lanes with readlane / writelane only.
template <int NX, int T>
struct LURows {
  static constexpr int NW = (NX + 3) / 4;
  double dinv;  // 1 / pivot of the row this lane holds
  int pos;      // elimination step at which this lane's row was the pivot

  // Factorises the rows a[] (lane i: row i) and stores the factors in S
  // (row i, column j at lu[j][i]: the multipliers left of the row's pivot
  // step, the row of U from it on).  Returns false for a zero or non-finite
  // pivot (recoverable setup failure).
  template <class L>
  __device__ __forceinline__ bool factor(const Team<T>& tm, double (&a)[NX], L& S) {
    const bool act = tm.lane < NX;
    pos = act ? -1 : NX;
    dinv = 0.0;
    bool ok = true;
    unsigned int pw = 0;  // pivots of the current group of four steps
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      if (!ok) continue;  // (team-uniform; keeps the loop fully unrolled)
      double v = (pos < 0) ? fabs(a[k]) : -1.0;
      if (v != v) v = INFINITY;  // a NaN entry fails the factorisation
      int p;
      const double mv = tm.argmax(v, p);
      if (!(mv > 0.0) || !(mv < INFINITY)) {
        ok = false;
        continue;
      }
      pw |= (unsigned int)p << (8 * (k & 3));
      if ((k & 3) == 3 || k == NX - 1) {
        if (tm.lane == 0) S.permw[k >> 2] = pw;
        pw = 0;
      }
      const double inv = 1.0 / tm.bcast(a[k], p);
      if (tm.lane == p) { pos = k; dinv = inv; }
      const bool upd = pos < 0;
      const double l = a[k] * inv;
      if (upd) a[k] = l;
#pragma unroll
      for (int j = k + 1; j < NX; ++j) {
        const double pj = tm.bcast(a[j], p);
        if (upd) a[j] = fma(-pj, l, a[j]);
      }
    }
    if (act) {
#pragma unroll
      for (int j = 0; j < NX; ++j) S.lu[j][tm.lane] = a[j];
    }
    return ok;
  }

  // Solves (LU) x = b; b is this lane's component of the right-hand side
  // (row order) and the result is x_lane.  Forward substitution: step k reads
  // y_k = b from the k-th pivot lane, and every lane subtracts its entry of
  // column k times y_k; lanes whose row was pivoted earlier collect
  // meaningless updates, so y_k is also kept in LDS (wsol[k], stored by every
  // lane with the same value, which costs less than a lane test), from where
  // every lane restarts the back substitution with y of its own pivot step.
  // Back substitution likewise: x_k = b / pivot on lane p_k, kept in wsol[k],
  // and x_lane is read from there at the end.
  template <class L>
  __device__ __forceinline__ double solve(double b, const Team<T>& tm, L& S) const {
    const bool act = tm.lane < NX;
    const int cl = act ? tm.lane : NX - 1;
    unsigned int pw[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) pw[w] = __builtin_amdgcn_readfirstlane(S.permw[w]);
    double r[NX];
#pragma unroll
    for (int k = 0; k < NX; ++k) r[k] = S.lu[k][cl];
#pragma unroll
    for (int k = 0; k < NX; ++k) {
      const int pk = (pw[k >> 2] >> (8 * (k & 3))) & 0xFF;
      const double yk = tm.bcast(b, pk);
      S.wsol[k] = yk;
      b = fma(-r[k], yk, b);
    }
    team_sync();
    b = S.wsol[act ? pos : 0];
    team_sync();
#pragma unroll
    for (int k = NX - 1; k >= 0; --k) {
      const int pk = (pw[k >> 2] >> (8 * (k & 3))) & 0xFF;
      const double xk = tm.bcast(b * dinv, pk);
      S.wsol[k] = xk;
      b = fma(-r[k], xk, b);
    }
    team_sync();
    const double x = S.wsol[cl];
    team_sync();
    return x;
  }
};

// Blocked row-per-lane dense LU of an NX x NX matrix (64 < NX <= T) for a
// multi-wavefront team.  The matrix lives in global memory, column-major
// (row i, column j at A[j * NX + i], so the lanes of a column access are
// consecutive), and is factorised in place; lane i owns row i.  Partial
// pivoting as in LURows (largest magnitude, ties to the smaller row, rows
// never move: perm[k] = row chosen at step k), in panels of NB columns
// (right-looking): the panel is eliminated in registers; per column, each
// wavefront's best candidate row publishes its magnitude, index and panel
// values through LDS in one exchange (double-buffered, one barrier).  The
// pivot rows' trailing entries become U12 = L11^-1 A12 (one lane per
// column), and the rows not yet chosen take the rank-NB update
// A22 -= L21 U12 with U12 staged in LDS.  Global loads are issued in
// batches ahead of the arithmetic and the stores, so one memory latency
// covers a batch.  The diagonal 16 x 16 blocks of the factors stay in LDS
// for the solves, which run block by block: the NB pivot values of a block
// are published through LDS (double-buffered, one barrier per block), every
// lane solves the small triangle, and the other rows update with their NB
// factors of the block.
template <int NX, int T>
struct LUBig {
  static_assert(NX > 64 && NX <= T, "LUBig: 64 < NX <= T");
  static constexpr int NB = 16;  // panel width / solve block
  static constexpr int CB = 8;   // trailing columns per batch of loads
  static constexpr int W = T / 64;
  static constexpr int NBLK = (NX + NB - 1) / NB;
  struct Shared {
    double u[NB][T + CB];       // U12 rows of the panel (column k1 + j at index j)
    double diag[NBLK][NB][NB];  // factors of each panel's pivot rows inside the panel
                                // (row c: pivot of step k0 + c; L left of c, U from c)
    double dinv[NBLK][NB];      // inverse pivots
    double cand[2][W][NB + 1];  // per wavefront: best |value|, then that row's panel values
    int cidx[2][W];             // per wavefront: row of the best candidate
    double y[2][NB];            // solve: values of the block's pivot rows
    int perm[NX];               // perm[k] = row chosen at step k
  };
  // per-lane results of a factorisation
  struct Fact {
    int ok, pos, src;
    double dinv;
  };
  static __device__ __forceinline__ Shared& shared() {
    __shared__ Shared sh;
    return sh;
  }

  double dinv;  // 1 / pivot of the row this lane holds
  int pos;      // elimination step at which this lane's row was the pivot
  int src;      // lane that holds pivot row number `lane`

  // The factorisation and the solve are separate (not inlined) functions
  // that take and return everything by value: their loops then get the
  // register file to themselves instead of competing with the integrator's
  // live state, which is saved only around the call.
  template <class L>
  __device__ __forceinline__ bool factor(const Team<T>&, double* A, L&) {
    const Fact f = factor_impl(A);
    pos = f.pos;
    src = f.src;
    dinv = f.dinv;
    return f.ok != 0;
  }
  template <class L>
  __device__ __forceinline__ double solve(double b, const Team<T>&, const double* A,
                                          const L&) const {
    return solve_impl(b, A, pos, src, dinv);
  }

  // Factorises A in place; ok = 0 for a zero or non-finite pivot.
  static __device__ __noinline__ Fact factor_impl(double* A) {
    Shared& sh = shared();
    const int i = threadIdx.x;
    const int wave = i >> 6;
    const bool act = i < NX;
    int pos = act ? -1 : NX;
    int src = 0;
    double dinv = 0.0;
    bool ok = true;
    for (int k0 = 0; k0 < NX && ok; k0 += NB) {
      const int kb = (NX - k0) < NB ? (NX - k0) : NB;
      double pa[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) pa[c] = (act && c < kb) ? A[(size_t)(k0 + c) * NX + i] : 0.0;
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        if (c >= kb || !ok) continue;  // team-uniform
        const int k = k0 + c;
        const int buf = k & 1;
        double v = (act && pos < 0) ? fabs(pa[c]) : -1.0;
        if (v != v) v = INFINITY;  // a NaN entry fails the factorisation
        int wi;
        const double wv = Wave64::argmax(v, i, wi);
        if (i == wi) {
          sh.cand[buf][wave][0] = wv;
          sh.cidx[buf][wave] = wi;
#pragma unroll
          for (int j = c; j < NB; ++j) sh.cand[buf][wave][1 + j] = pa[j];
        }
        lds_barrier();
        // wavefronts in order, ties to the smaller row: team-uniform winner
        double mv = sh.cand[buf][0][0];
        int p = sh.cidx[buf][0], pw = 0;
#pragma unroll
        for (int w = 1; w < W; ++w) {
          const double rv = sh.cand[buf][w][0];
          const int ri = sh.cidx[buf][w];
          if (rv > mv || (rv == mv && ri < p)) { mv = rv; p = ri; pw = w; }
        }
        if (!(mv > 0.0) || !(mv < INFINITY)) {
          ok = false;
          continue;
        }
        double pr[NB];
#pragma unroll
        for (int j = c; j < NB; ++j) pr[j] = sh.cand[buf][pw][1 + j];
        if (i == 0) sh.perm[k] = p;
        if (i == k) src = p;
        const double inv = 1.0 / pr[c];
        if (i == p) { pos = k; dinv = inv; }
        if (act && pos < 0) {
          const double l = pa[c] * inv;
          pa[c] = l;
#pragma unroll
          for (int j = c + 1; j < NB; ++j) pa[j] = fma(-pr[j], l, pa[j]);
        }
      }
      if (!ok) break;  // team-uniform
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (act && c < kb) A[(size_t)(k0 + c) * NX + i] = pa[c];
      const int k1 = k0 + kb;
      // factors of the panel's pivot rows inside the panel, kept for the solves
      if (act && pos >= k0 && pos < k1) {
        const int c = pos - k0;
#pragma unroll
        for (int cc = 0; cc < NB; ++cc) sh.diag[k0 / NB][c][cc] = pa[cc];
        sh.dinv[k0 / NB][c] = dinv;
      }
      if (k1 >= NX) break;
      __syncthreads();  // the panel's rows in global memory are read by other lanes next
      // U12: lane j finishes the pivot rows' entries of trailing column k1 + j
      const int col = k1 + i;
      if (col < NX) {
        int prr[NB];
        double u[NB];
#pragma unroll
        for (int c = 0; c < NB; ++c) prr[c] = (c < kb) ? sh.perm[k0 + c] : 0;
#pragma unroll
        for (int c = 0; c < NB; ++c) u[c] = (c < kb) ? A[(size_t)col * NX + prr[c]] : 0.0;
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < kb) {
#pragma unroll
            for (int cc = 0; cc < c; ++cc) u[c] = fma(-sh.diag[k0 / NB][c][cc], u[cc], u[c]);
          }
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < kb) A[(size_t)col * NX + prr[c]] = u[c];
#pragma unroll
        for (int c = 0; c < NB; ++c) sh.u[c][i] = u[c];
      }
      lds_barrier();  // U12 rows in LDS; the rows updated next are the lanes' own
      // A22 -= L21 U12 on the rows not chosen yet, CB columns per batch
      if (act && pos < 0) {
        const int nc = NX - k1;
        for (int j0 = 0; j0 < nc; j0 += CB) {
          double v[CB];
#pragma unroll
          for (int q = 0; q < CB; ++q)
            v[q] = (j0 + q < nc) ? A[(size_t)(k1 + j0 + q) * NX + i] : 0.0;
#pragma unroll
          for (int c = 0; c < NB; ++c) {
            const double l = pa[c];
#pragma unroll
            for (int q = 0; q < CB; ++q) v[q] = fma(-l, sh.u[c][j0 + q], v[q]);
          }
#pragma unroll
          for (int q = 0; q < CB; ++q)
            if (j0 + q < nc) A[(size_t)(k1 + j0 + q) * NX + i] = v[q];
        }
      }
      lds_barrier();
    }
    // U12 entries were written by other lanes than the rows' owners, which
    // read them in the solves
    __syncthreads();
    return Fact{ok ? 1 : 0, pos, src, dinv};
  }

  // Solves (LU) x = b with the factors in A; b is this lane's component.
  // The diagonal blocks of the factors come from LDS (kept by factor()), so
  // a block's critical path is one LDS exchange of the pivot values; the
  // other rows' loads of their own factors overlap it.
  static __device__ __noinline__ double solve_impl(double b, const double* A, int pos, int src,
                                                   double dinv) {
    Shared& sh = shared();
    const Team<T> tm;
    const int i = tm.lane;
    const bool act = i < NX;
    int blk = 0;  // running block count: LDS buffer parity
    // forward substitution L y = P b, NB pivots at a time
    for (int kb0 = 0; kb0 < NBLK; ++kb0, ++blk) {
      const int k0 = kb0 * NB;
      const int kb = (NX - k0) < NB ? (NX - k0) : NB;
      const int buf = blk & 1;
      const bool piv = act && pos >= k0 && pos < k0 + kb;
      const bool upd = act && pos >= k0 + kb;
      double f[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) f[c] = (upd && c < kb) ? A[(size_t)(k0 + c) * NX + i] : 0.0;
      if (piv) sh.y[buf][pos - k0] = b;
      lds_barrier();
      double y[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) {
        double v = 0.0;
        if (c < kb) {
          v = sh.y[buf][c];
#pragma unroll
          for (int cc = 0; cc < c; ++cc) v = fma(-sh.diag[kb0][c][cc], y[cc], v);
        }
        y[c] = v;
      }
      if (piv) {
        b = sel(y, pos - k0);
      } else if (upd) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < kb) b = fma(-f[c], y[c], b);
      }
    }
    // back substitution U x = y, blocks from the last one
    for (int kb0 = NBLK - 1; kb0 >= 0; --kb0, ++blk) {
      const int k0 = kb0 * NB;
      const int kb = (NX - k0) < NB ? (NX - k0) : NB;
      const int buf = blk & 1;
      const bool piv = act && pos >= k0 && pos < k0 + kb;
      const bool upd = act && pos < k0;
      double f[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) f[c] = (upd && c < kb) ? A[(size_t)(k0 + c) * NX + i] : 0.0;
      if (piv) sh.y[buf][pos - k0] = b;
      lds_barrier();
      double z[NB];
#pragma unroll
      for (int c = NB - 1; c >= 0; --c) {
        double v = 0.0;
        if (c < kb) {
          v = sh.y[buf][c];
#pragma unroll
          for (int cc = c + 1; cc < NB; ++cc)
            if (cc < kb) v = fma(-sh.diag[kb0][c][cc], z[cc], v);
          v *= sh.dinv[kb0][c];
        }
        z[c] = v;
      }
      if (piv) {
        b = sel(z, pos - k0);
      } else if (upd) {
#pragma unroll
        for (int c = 0; c < NB; ++c)
          if (c < kb) b = fma(-f[c], z[c], b);
      }
    }
    return tm.gather(b, act ? src : 0);
  }
};

constexpr int TRACE_REC = 8;

// Optional cycle profile (diagnostic builds compiled with -DAMX_PROFILE):
// shader-clock cycles spent in each part of the integrator, accumulated per
// instance.  Product builds compile the hooks away.
enum : int {
  PF_STEP, PF_PREDICT, PF_COEFF, PF_NEWTON, PF_RHS, PF_LSETUP, PF_SOLVE, PF_ERRTEST,
  PF_QUAD, PF_COMPLETE, PF_PREPARE, PF_LOOP, PF_REINIT, PF_N = 16
};
#ifdef AMX_PROFILE
#define AMX_PF_BEGIN(v) const unsigned long long v = __builtin_readcyclecounter()
#define AMX_PF_END(v, k) prof[k] += __builtin_readcyclecounter() - (v)
#else
#define AMX_PF_BEGIN(v)
#define AMX_PF_END(v, k)
#endif

struct SolverParams {
  double rtol, atol;
  double rtolQ, atolQ;
  double rtolS, atolS;
  int maxsteps;
  int errconQ;
  int stldet;                   // BDF stability-limit detection on (CVodeSetStabLimDet)
  unsigned long long deadline;  // wall_clock64() limit, 0 = none
};


// Optional problem properties (false when the problem does not declare them):
//   P::FORCE_SETUP  the integrator honours force_setup (a fresh Jacobian and
//                   Newton matrix at the next nonlinear solve)
//   P::CK_SAVE      after every accepted step the integrator writes its state
//                   to pb.ck_rec + pb.ck_due * CK_LEN when pb.ck_due >= 0
template <class P, class = void>
struct force_setup_of : std::false_type {};
template <class P>
struct force_setup_of<P, std::void_t<decltype(P::FORCE_SETUP)>>
    : std::bool_constant<P::FORCE_SETUP> {};
template <class P, class = void>
struct ck_save_of : std::false_type {};
template <class P>
struct ck_save_of<P, std::void_t<decltype(P::CK_SAVE)>> : std::bool_constant<P::CK_SAVE> {};

// scalars of an integrator checkpoint after the Nordsieck array (save_state)
constexpr int CK_SCAL = 72;

// BDF integrator over a Problem P.  P provides:
//   NX, T, NQ (quadratures, 0 if none), NS (forward sensitivities, 0 if none)
//   double rhs(double t, double y)                    team-collective, returns f_lane
//   void   quad_rhs(double t, double y, double* fq)   fq[EPQ] (if NQ > 0)
//   void   jac(double t, double y)                    assembles J into its LDS
//   void   build_M(double gamma, double* a)           a[j] = (I - gamma*J)[lane][j]
//   bool   on_step(t, zn0, zn1, hscale, q_used)       after every accepted step
//   sensitivity hooks (if NS > 0): see sens_* in the problem
template <class P>
struct BDF {
  static constexpr int NX = P::NX;
  static constexpr int T = P::T;
  static constexpr int NQ = P::NQ;
  static constexpr int EPQ = (NQ > 0) ? (NQ + T - 1) / T : 1;
  static constexpr bool QUAD = NQ > 0;
  static constexpr bool SENS = P::NS > 0;
  static constexpr bool FORCE = force_setup_of<P>::value;
  static constexpr bool CK_SAVE = ck_save_of<P>::value;
  // doubles of a checkpoint record: zn[LMAX][NX], then CK_SCAL scalars
  static constexpr int CK_LEN = LMAX * NX + CK_SCAL;
  using L = BdfLds<NX, NQ>;

  P& pb;
  L& S;
  Team<T> tm;
  SolverParams sp;
  bool active;  // lane < NX
  int nq;       // quadratures integrated (<= NQ; the problem's parameter list)

  // per-lane work values
  double ewt, acor, y, ftemp;
  double ewtQ[EPQ], acorQ[EPQ], yQ[EPQ];
  // LU in LDS and registers for one-wavefront teams, in global memory
  // (the problem's per-instance matrix pb.mat) for multi-wavefront teams
  static constexpr bool GLU = T > 64;
  std::conditional_t<GLU, LUBig<NX, T>, LURows<NX, T>> lu;

  // team-uniform scalars
  double tn, h, hprime, hscale, eta, etamax, etaq, etaqm1, etaqp1, gamma,
      gammap, gamrat, rl1, crate, crateS, delp, acnrm, acnrmS, saved_tq5, hu,
      tolsf, tstop, dsm_last, dsm_test;
  int q, qprime, L_, qwait, nst, nstlp, nstlj, nscon, convfail;
  bool jcur, tstopset, at_tstop;
  bool force_setup;  // FORCE problems: fresh Jacobian and Newton matrix at the next solve
  // values returned by advance(): this lane's state and quadratures
  double yout, yQout[EPQ];
  // counters
  int nfe, nsetups, nje, nni, netf, ncfn, nfQe, nniS, nor;
  // optional step trace (team lane 0 writes records of TRACE_REC doubles)
  double* trace;
  int trace_cap, trace_n;
#ifdef AMX_PROFILE
  unsigned long long prof[PF_N];
#endif

  __device__ __forceinline__ BDF(P& p, L& lds, const SolverParams& s) : pb(p), S(lds), sp(s) {
    active = tm.lane < NX;
    force_setup = false;
    if constexpr (QUAD) nq = p.nq;
    else nq = 0;
    trace = nullptr;
    trace_cap = 0;
    trace_n = 0;
#ifdef AMX_PROFILE
    for (int k = 0; k < PF_N; ++k) prof[k] = 0;
#endif
  }

  // ---------------- LDS accessors ----------------
  // Nordsieck column j of this lane's state (0 on lanes >= NX).
  __device__ __forceinline__ double Z(int j) const {
    const double v = S.zn[j][active ? tm.lane : 0];
    return active ? v : 0.0;
  }
  __device__ __forceinline__ void setZ(int j, double v) {
    if (active) S.zn[j][tm.lane] = v;
  }
  __device__ __forceinline__ int qslot(int e) const { return tm.lane + e * T; }
  __device__ __forceinline__ double ZQ(int j, int e) const {
    const int g = qslot(e);
    const double v = S.znQ[j][g < NQ ? g : 0];
    return g < NQ ? v : 0.0;
  }
  __device__ __forceinline__ void setZQ(int j, int e, double v) {
    const int g = qslot(e);
    if (g < NQ) S.znQ[j][g] = v;
  }

  // event: 1 accepted step, 2 nonlinear-solver failure, 3 error-test failure
  __device__ __forceinline__ void record(int event, double v) {
    if (trace == nullptr || trace_n >= trace_cap) return;
    if (tm.lane == 0) {
      double* r = trace + (size_t)trace_n * TRACE_REC;
      r[0] = event; r[1] = tn; r[2] = h; r[3] = q; r[4] = v;
      r[5] = nni; r[6] = nsetups; r[7] = nje;
    }
    trace_n++;
  }

  // ---------------- vector helpers ----------------
  __device__ __forceinline__ double wrms(double v, double w) const {
    const double e = active ? v * w : 0.0;
    return sqrt(tm.sum(e * e) / NX);
  }
  __device__ __forceinline__ double wrmsQ(const double (&v)[EPQ], const double (&w)[EPQ]) const {
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < EPQ; ++k) {
      const double e = (qslot(k) < nq) ? v[k] * w[k] : 0.0;
      acc += e * e;
    }
    return sqrt(tm.sum(acc) / (nq > 0 ? nq : 1));
  }
  // true when any lane holds a non-finite value (NaN-safe, team-uniform)
  __device__ __forceinline__ bool any_nonfinite(double v) const {
    return tm.any(active && !isfinite(v));
  }
  __device__ __forceinline__ void set_ewt() {
    ewt = active ? 1.0 / (sp.rtol * fabs(Z(0)) + sp.atol) : 0.0;
  }
  __device__ __forceinline__ void set_ewtQ() {
#pragma unroll
    for (int k = 0; k < EPQ; ++k)
      ewtQ[k] = (qslot(k) < nq) ? 1.0 / (sp.rtolQ * fabs(ZQ(0, k)) + sp.atolQ) : 0.0;
  }
  __device__ __forceinline__ bool quad_errcon() const { return QUAD && sp.errconQ; }
  __device__ __forceinline__ void znQ0(double (&v)[EPQ]) const {
#pragma unroll
    for (int e = 0; e < EPQ; ++e) v[e] = ZQ(0, e);
  }

  // ---------------- initialisation ----------------
  // CVodeInit state (cvodes.c:677-886) for y0 on this lane, quadratures q0.
  __device__ __forceinline__ void init(double t0, double y0, const double* q0) {
    tn = t0;
    team_sync();
    for (int j = 0; j < LMAX; ++j) setZ(j, 0.0);
    setZ(0, y0);
    if constexpr (QUAD) {
      for (int j = 0; j < LMAX; ++j)
#pragma unroll
        for (int e = 0; e < EPQ; ++e) setZQ(j, e, 0.0);
#pragma unroll
      for (int e = 0; e < EPQ; ++e) setZQ(0, e, q0 ? q0[e] : 0.0);
    }
    if (tm.lane == 0) {
      for (int j = 0; j <= LMAX; ++j) S.tau[j] = 0.0;
      for (int j = 0; j < LMAX; ++j) S.l[j] = 0.0;
      for (int j = 0; j < 6; ++j) S.tq[j] = 0.0;
      for (int i = 0; i < 5; ++i)
        for (int k = 0; k < 3; ++k) S.ssdat[i][k] = 0.0;
    }
    team_sync();
    q = 1; L_ = 2; qwait = L_; qprime = 1; etamax = ETA_MAX_FS;
    hu = 0.0; tolsf = 1.0; h = 0.0; hprime = 0.0; hscale = 0.0; eta = 1.0;
    nst = 0; nstlp = 0; nstlj = 0; nscon = 0; saved_tq5 = 0.0;
    gammap = 0.0; gamma = 0.0; gamrat = 1.0; crate = 1.0; crateS = 1.0;
    delp = 0.0; acnrm = 0.0; acnrmS = 0.0; convfail = NO_FAILURES;
    jcur = false; tstopset = false; at_tstop = false; tstop = 0.0;
    etaq = etaqm1 = etaqp1 = 0.0;
    yout = active ? y0 : 0.0;
    nfe = nsetups = nje = nni = netf = ncfn = nfQe = nniS = nor = 0;
    dsm_last = 0.0; dsm_test = 0.0;
    ewt = 0.0; acor = 0.0; y = active ? y0 : 0.0; ftemp = 0.0;
#pragma unroll
    for (int e = 0; e < EPQ; ++e) ewtQ[e] = acorQ[e] = yQ[e] = yQout[e] = 0.0;
  }

  // ---------------- Nordsieck array manipulation ----------------
  // zn[j-1] += zn[j] for k = 1..q, j = q..k (cvPredict)
  __device__ __forceinline__ void pascal(double sign) {
    if (active) {
      double z[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? S.zn[j][tm.lane] : 0.0;
#pragma unroll
      for (int k = 1; k <= QMAX; ++k)
#pragma unroll
        for (int j = QMAX; j >= k; --j)
          if (k <= q && j <= q) z[j - 1] = fma(sign, z[j], z[j - 1]);
#pragma unroll
      for (int j = 0; j < QMAX; ++j)
        if (j < q) S.zn[j][tm.lane] = z[j];
    }
    if constexpr (QUAD) {
#pragma unroll
      for (int e = 0; e < EPQ; ++e) {
        const int g = qslot(e);
        if (g < NQ) {
          double z[LMAX];
#pragma unroll
          for (int j = 0; j < LMAX; ++j) z[j] = (j <= q) ? S.znQ[j][g] : 0.0;
#pragma unroll
          for (int k = 1; k <= QMAX; ++k)
#pragma unroll
            for (int j = QMAX; j >= k; --j)
              if (k <= q && j <= q) z[j - 1] = fma(sign, z[j], z[j - 1]);
#pragma unroll
          for (int j = 0; j < QMAX; ++j)
            if (j < q) S.znQ[j][g] = z[j];
        }
      }
    }
  }

  __device__ __forceinline__ void predict() {
    tn += h;
    if (tstopset && (tn - tstop) * h > 0.0) tn = tstop;
    pascal(1.0);
    if constexpr (SENS) pb.sens_predict(q);
  }

  __device__ __forceinline__ void restore(double saved_t) {
    tn = saved_t;
    pascal(-1.0);
    if constexpr (SENS) pb.sens_restore(q);
  }

  __device__ __forceinline__ void rescale() {
    double c = eta;
#pragma unroll
    for (int j = 1; j <= QMAX; ++j) {
      if (j <= q) {
        setZ(j, Z(j) * c);
        if constexpr (QUAD)
#pragma unroll
          for (int e = 0; e < EPQ; ++e) setZQ(j, e, ZQ(j, e) * c);
      }
      c *= eta;
    }
    if constexpr (SENS) pb.sens_rescale(q, eta);
    h = hscale * eta;
    hscale = h;
    nscon = 0;
  }

  // cvIncreaseBDF: zn[q+1] = A1 zn[QMAX] (saved acor), zn[j] += l[j] zn[q+1]
  __device__ __forceinline__ void increase_bdf() {
    double ll[LMAX];
#pragma unroll
    for (int i = 0; i < LMAX; ++i) ll[i] = 0.0;
    ll[2] = 1.0;
    double alpha1 = 1.0, prod = 1.0, xiold = 1.0, alpha0 = -1.0, hsum = hscale;
#pragma unroll
    for (int j = 1; j < QMAX; ++j) {
      if (j < q) {
        hsum += S.tau[j + 1];
        const double xi = hsum / hscale;
        prod *= xi;
        alpha0 -= 1.0 / (j + 1);
        alpha1 += 1.0 / xi;
#pragma unroll
        for (int i = j + 2; i >= 2; --i)
          if (i < LMAX) ll[i] = ll[i] * xiold + ll[i - 1];
        xiold = xi;
      }
    }
    const double A1 = (-alpha0 - alpha1) / prod;
    const double znL = A1 * Z(QMAX);
#pragma unroll
    for (int j = 2; j <= QMAX; ++j)
      if (j <= q) setZ(j, fma(ll[j], znL, Z(j)));
    if (q + 1 <= QMAX) setZ(q + 1, znL);
    if constexpr (QUAD) {
#pragma unroll
      for (int e = 0; e < EPQ; ++e) {
        const double zQL = A1 * ZQ(QMAX, e);
#pragma unroll
        for (int j = 2; j <= QMAX; ++j)
          if (j <= q) setZQ(j, e, fma(ll[j], zQL, ZQ(j, e)));
        if (q + 1 <= QMAX) setZQ(q + 1, e, zQL);
      }
    }
    if constexpr (SENS) pb.sens_increase(q, A1, ll);
  }

  // cvDecreaseBDF
  __device__ __forceinline__ void decrease_bdf() {
    double ll[LMAX];
#pragma unroll
    for (int i = 0; i < LMAX; ++i) ll[i] = 0.0;
    ll[2] = 1.0;
    double hsum = 0.0;
#pragma unroll
    for (int j = 1; j <= QMAX - 2; ++j) {
      if (j <= q - 2) {
        hsum += S.tau[j];
        const double xi = hsum / hscale;
#pragma unroll
        for (int i = j + 2; i >= 2; --i) ll[i] = ll[i] * xi + ll[i - 1];
      }
    }
    if (q > 2) {
      const double znq = Z(q);
#pragma unroll
      for (int j = 2; j < QMAX; ++j)
        if (j < q) setZ(j, fma(-ll[j], znq, Z(j)));
      if constexpr (QUAD) {
#pragma unroll
        for (int e = 0; e < EPQ; ++e) {
          const double zq = ZQ(q, e);
#pragma unroll
          for (int j = 2; j < QMAX; ++j)
            if (j < q) setZQ(j, e, fma(-ll[j], zq, ZQ(j, e)));
        }
      }
      if constexpr (SENS) pb.sens_decrease(q, ll);
    }
  }

  __device__ __forceinline__ void adjust_order(int deltaq) {
    if (q == 2 && deltaq != 1) return;
    if (deltaq == 1) increase_bdf();
    else if (deltaq == -1) decrease_bdf();
  }

  __device__ __forceinline__ void adjust_params() {
    if (qprime != q) {
      adjust_order(qprime - q);
      q = qprime;
      L_ = q + 1;
      qwait = L_;
    }
    rescale();
  }

  // ---------------- method coefficients: cvSetBDF / cvSetTqBDF ----------------
  __device__ __forceinline__ void set_coeffs() {
    double l[LMAX], tq[6];
    double xi_inv = 1.0, xistar_inv = 1.0, alpha0 = -1.0, alpha0_hat = -1.0;
    double hsum = h;
#pragma unroll
    for (int i = 0; i < LMAX; ++i) l[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; ++i) tq[i] = S.tq[i];
    l[0] = l[1] = 1.0;
    double lq = 1.0;
    if (q > 1) {
#pragma unroll
      for (int j = 2; j < QMAX; ++j) {
        if (j < q) {
          hsum += S.tau[j - 1];
          xi_inv = h / hsum;
          alpha0 -= 1.0 / j;
#pragma unroll
          for (int i = j; i >= 1; --i) l[i] += l[i - 1] * xi_inv;
        }
      }
      alpha0 -= 1.0 / q;
      xistar_inv = -l[1] - alpha0;
      hsum += S.tau[q - 1];
      xi_inv = h / hsum;
      alpha0_hat = -l[1] - xi_inv;
#pragma unroll
      for (int i = QMAX; i >= 1; --i)
        if (i <= q) l[i] += l[i - 1] * xistar_inv;
#pragma unroll
      for (int i = 2; i <= QMAX; ++i)
        if (i == q) lq = l[i];
    }
    const double A1 = 1.0 - alpha0_hat + alpha0;
    const double A2 = 1.0 + q * A1;
    tq[2] = fabs(A1 / (alpha0 * A2));
    tq[5] = fabs(A2 * xistar_inv / (lq * xi_inv));
    if (qwait == 1) {
      if (q > 1) {
        const double C = xistar_inv / lq;
        const double A3 = alpha0 + 1.0 / q;
        const double A4 = alpha0_hat + xi_inv;
        const double Cpinv = (1.0 - A4 + A3) / A3;
        tq[1] = fabs(C * Cpinv);
      } else {
        tq[1] = 1.0;
      }
      hsum += S.tau[q];
      xi_inv = h / hsum;
      const double A5 = alpha0 - 1.0 / (q + 1);
      const double A6 = alpha0_hat - xi_inv;
      const double Cppinv = (1.0 - A6 + A5) / A2;
      tq[3] = fabs(Cppinv / (xi_inv * (q + 2) * A5));
    }
    tq[4] = CORTES / tq[2];
    team_sync();
    if (tm.lane == 0) {
#pragma unroll
      for (int i = 0; i < LMAX; ++i) S.l[i] = l[i];
#pragma unroll
      for (int i = 1; i < 6; ++i) S.tq[i] = tq[i];
    }
    team_sync();
    rl1 = 1.0 / l[1];
    gamma = h * rl1;
    if (nst == 0) gammap = gamma;
    gamrat = (nst > 0) ? gamma / gammap : 1.0;
  }

  // ---------------- linear system ----------------
  // cvLsSetup + cvNlsLSetup; returns NLS_OK or NLS_RECOVER (singular).
  __device__ __forceinline__ int lsetup(int cf) {
    AMX_PF_BEGIN(t0);
    const double dgamma = fabs(gamma / gammap - 1.0);
    const bool jbad = (nst == 0) || (nst >= nstlj + MSBJ) ||
                      (cf == FAIL_BAD_J && dgamma < DGMAX_JBAD) || (cf == FAIL_OTHER) ||
                      (FORCE && force_setup);
    if constexpr (FORCE) force_setup = false;
    if (jbad) {
      pb.jac(tn, y);
      jcur = true;
      nje++;
      nstlj = nst;
    } else {
      jcur = false;
    }
    bool ok;
    if constexpr (GLU) {
      pb.build_M_global(gamma);
      ok = lu.factor(tm, pb.mat, S);
    } else {
      double a[NX];
      pb.build_M(gamma, a);
      ok = lu.factor(tm, a, S);
    }
    nsetups++;
    gamrat = 1.0;
    gammap = gamma;
    crate = 1.0;
    crateS = 1.0;
    nstlp = nst;
    AMX_PF_END(t0, PF_LSETUP);
    return ok ? NLS_OK : NLS_RECOVER;
  }

  __device__ __forceinline__ double lsolve(double b) {
    AMX_PF_BEGIN(t0);
    double x;
    if constexpr (GLU) x = lu.solve(b, tm, pb.mat, S);
    else x = lu.solve(b, tm, S);
    if (gamrat != 1.0) x *= 2.0 / (1.0 + gamrat);
    AMX_PF_END(t0, PF_SOLVE);
    return x;
  }

  // ---------------- nonlinear solve (Newton) ----------------
  // cvNlsResidual: y = zn0 + acor, ftemp = f(tn, y), returns residual.
  __device__ __forceinline__ double residual() {
    y = Z(0) + acor;
    AMX_PF_BEGIN(t0);
    ftemp = pb.rhs(tn, y);
    AMX_PF_END(t0, PF_RHS);
    nfe++;
    return rl1 * Z(1) + acor - gamma * ftemp;
  }

  // SUNNonlinSolSolve_Newton with cvNlsConvTest; returns NLS_OK / NLS_RECOVER.
  // A non-finite residual or correction is a recoverable failure; the number
  // of fresh-Jacobian retries is bounded (CVODES needs at most two).
  __device__ __forceinline__ int newton(bool callSetup) {
    bool jbad = false;
    for (int attempt = 0; attempt < 3; ++attempt) {
      int m = 0;
      double res = residual();
      int ret = NLS_OK;
      if (any_nonfinite(res)) ret = NLS_RECOVER;
      if (ret == NLS_OK && callSetup) {
        if (jbad) convfail = FAIL_BAD_J;
        ret = lsetup(convfail);
      }
      if (ret == NLS_OK) {
        for (int it = 0; it < NLS_MAXCOR; ++it) {
          nni++;
          double delta = lsolve(-res);
          if (!active) delta = 0.0;
          acor += delta;
          const double del = wrms(delta, ewt);
          if (!(del == del)) { ret = NLS_RECOVER; break; }
          if (m > 0) crate = fmax(CRDOWN * crate, del / delp);
          const double dcon = del * fmin(1.0, crate) / S.tq[4];
          if (dcon <= 1.0) {
            acnrm = (m == 0) ? del : wrms(acor, ewt);
            jcur = false;
            return NLS_OK;
          }
          if (m >= 1 && del > RDIV * delp) { ret = NLS_RECOVER; break; }
          delp = del;
          m++;
          if (m >= NLS_MAXCOR) { ret = NLS_RECOVER; break; }
          res = residual();
          if (any_nonfinite(res)) { ret = NLS_RECOVER; break; }
        }
      }
      if (ret == NLS_RECOVER && !jcur) {
        callSetup = true;
        jbad = true;
        acor = 0.0;
        continue;
      }
      return ret;
    }
    return NLS_RECOVER;
  }

  __device__ __forceinline__ int nls(int nflag) {
    convfail = (nflag == FIRST_CALL || nflag == PREV_ERR_FAIL) ? NO_FAILURES : FAIL_OTHER;
    const bool callSetup = (nflag == PREV_CONV_FAIL) || (nflag == PREV_ERR_FAIL) ||
                           (nst == 0) || (nst >= nstlp + MSBP) ||
                           (fabs(gamrat - 1.0) > DGMAX_LSETUP) || (FORCE && force_setup);
    acor = 0.0;
    const int flag = newton(callSetup);
    if (flag != NLS_OK) return flag;
    y = Z(0) + acor;
    jcur = false;
    return NLS_OK;
  }

  // cvHandleNFlag
  __device__ __forceinline__ int handle_nflag(int& nflag, double saved_t, int& ncf) {
    if (nflag == NLS_OK) return DO_ERROR_TEST;
    record(2, (double)nflag);
    ncfn++;
    restore(saved_t);
    ncf++;
    etamax = 1.0;
    if (ncf == MXNCF) return AMX_CONV_FAILURE;
    eta = ETA_CF;
    nflag = PREV_CONV_FAIL;
    rescale();
    return PREDICT_AGAIN;
  }

  // cvDoErrorTest (state part); the order-1 restart branch re-evaluates the
  // quadrature / sensitivity derivatives.
  __device__ __forceinline__ int error_test(int& nflag, double saved_t, double acor_nrm, int& nef) {
    const double dsm = acor_nrm * S.tq[2];
    dsm_test = dsm;
    if (dsm <= 1.0) return AMX_SUCCESS;
    record(3, dsm);
    nef++;
    netf++;
    nflag = PREV_ERR_FAIL;
    restore(saved_t);
    if (nef == MXNEF) return AMX_ERR_FAILURE;
    etamax = 1.0;
    if (nef <= MXNEF1) {
      eta = 1.0 / (powr(BIAS2 * dsm, 1.0 / L_) + ADDON);
      eta = fmax(ETA_MIN_EF, eta);
      if (nef >= SMALL_NEF) eta = fmin(eta, ETA_MAX_EF);
      rescale();
      return TRY_AGAIN;
    }
    if (q > 1) {
      eta = ETA_MIN_EF;
      adjust_order(-1);
      L_ = q;
      q--;
      qwait = L_;
      rescale();
      return TRY_AGAIN;
    }
    eta = ETA_MIN_EF;
    h *= eta;
    hscale = h;
    qwait = LONG_WAIT;
    nscon = 0;
    const double f0 = pb.rhs(tn, Z(0));
    nfe++;
    setZ(1, h * f0);
    if constexpr (QUAD) {
      double fq[EPQ];
      pb.quad_rhs(tn, Z(0), fq);
      nfQe++;
#pragma unroll
      for (int e = 0; e < EPQ; ++e) setZQ(1, e, h * fq[e]);
    }
    if constexpr (SENS) pb.sens_restart_q1(tn, Z(0), f0, h);
    return TRY_AGAIN;
  }

  // cvQuadNls
  __device__ __forceinline__ void quad_correct() {
    double fq[EPQ];
    pb.quad_rhs(tn, y, fq);
    nfQe++;
#pragma unroll
    for (int e = 0; e < EPQ; ++e) {
      acorQ[e] = rl1 * (h * fq[e] - ZQ(1, e));
      yQ[e] = ZQ(0, e) + acorQ[e];
    }
  }

  // cvCompleteStep
  __device__ __forceinline__ void complete_step() {
    nst++;
    nscon++;
    hu = h;
    double tau[LMAX + 1];
#pragma unroll
    for (int i = 0; i <= LMAX; ++i) tau[i] = S.tau[i];
#pragma unroll
    for (int i = QMAX; i >= 2; --i)
      if (i <= q) tau[i] = tau[i - 1];
    if (q == 1 && nst > 1) tau[2] = tau[1];
    tau[1] = h;
    team_sync();
    if (tm.lane == 0) {
#pragma unroll
      for (int i = 0; i <= LMAX; ++i) S.tau[i] = tau[i];
    }
    team_sync();
    for (int j = 0; j <= q; ++j) {
      const double lj = S.l[j];
      setZ(j, fma(lj, acor, Z(j)));
      if constexpr (QUAD)
#pragma unroll
        for (int e = 0; e < EPQ; ++e) setZQ(j, e, fma(lj, acorQ[e], ZQ(j, e)));
    }
    if constexpr (SENS) {
      double l[LMAX];
#pragma unroll
      for (int j = 0; j < LMAX; ++j) l[j] = S.l[j];
      pb.sens_complete(q, l);
    }
    qwait--;
    if (qwait == 1 && q != QMAX) {
      setZ(QMAX, acor);
      if constexpr (QUAD)
#pragma unroll
        for (int e = 0; e < EPQ; ++e) setZQ(QMAX, e, acorQ[e]);
      if constexpr (SENS) pb.sens_save_acor();
      saved_tq5 = S.tq[5];
    }
  }

  __device__ __forceinline__ void set_eta() {
    if (eta > 0.0 && eta < ETA_MAX_FX) {
      eta = 1.0;
      hprime = h;
    } else {
      if (eta >= ETA_MAX_FX) {
        eta = fmin(eta, etamax);
      } else {
        eta = fmax(eta, ETA_MIN);
      }
      hprime = h * eta;
      if (qprime < q) nscon = 0;
    }
  }

  __device__ __forceinline__ double norm_with_extras(double nstate, const double (&vQ)[EPQ],
                                                     bool useQ, double sensnorm, bool useS) const {
    double r = nstate;
    if constexpr (QUAD)
      if (useQ) r = fmax(r, wrmsQ(vQ, ewtQ));
    if (useS) r = fmax(r, sensnorm);
    return r;
  }

  __device__ __forceinline__ double compute_etaqm1() {
    etaqm1 = 0.0;
    if (q > 1) {
      double zq[EPQ];
#pragma unroll
      for (int e = 0; e < EPQ; ++e) zq[e] = ZQ(q, e);
      double ddn = wrms(Z(q), ewt);
      double sn = 0.0;
      if constexpr (SENS) sn = pb.sens_norm_zn(q);
      ddn = norm_with_extras(ddn, zq, quad_errcon(), sn, SENS) * S.tq[1];
      etaqm1 = 1.0 / (powr(BIAS1 * ddn, 1.0 / q) + ADDON);
    }
    return etaqm1;
  }

  __device__ __forceinline__ double compute_etaqp1() {
    etaqp1 = 0.0;
    if (q != QMAX) {
      if (saved_tq5 == 0.0) return etaqp1;
      const double cquot = (S.tq[5] / saved_tq5) * powr(h / S.tau[2], (double)L_);
      const double tv = acor - cquot * Z(QMAX);
      double tQ[EPQ];
#pragma unroll
      for (int e = 0; e < EPQ; ++e) tQ[e] = acorQ[e] - cquot * ZQ(QMAX, e);
      double dup = wrms(tv, ewt);
      double sn = 0.0;
      if constexpr (SENS) sn = pb.sens_norm_qp1(cquot);
      dup = norm_with_extras(dup, tQ, quad_errcon(), sn, SENS) * S.tq[3];
      etaqp1 = 1.0 / (powr(BIAS3 * dup, 1.0 / (L_ + 1)) + ADDON);
    }
    return etaqp1;
  }

  __device__ __forceinline__ void choose_eta() {
    const double etam = fmax(etaqm1, fmax(etaq, etaqp1));
    if (etam > 0.0 && etam < ETA_MAX_FX) {
      eta = 1.0;
      qprime = q;
    } else if (etam == etaq) {
      eta = etaq;
      qprime = q;
    } else if (etam == etaqm1) {
      eta = etaqm1;
      qprime = q - 1;
    } else {
      eta = etaqp1;
      qprime = q + 1;
      setZ(QMAX, acor);
      if constexpr (QUAD)
        if (quad_errcon())
#pragma unroll
          for (int e = 0; e < EPQ; ++e) setZQ(QMAX, e, acorQ[e]);
      if constexpr (SENS) pb.sens_save_acor();
    }
  }

  __device__ __forceinline__ void prepare_next_step(double dsm) {
    if (etamax == 1.0) {
      qwait = max(qwait, 2);
      qprime = q;
      hprime = h;
      eta = 1.0;
      return;
    }
    etaq = 1.0 / (powr(BIAS2 * dsm, 1.0 / L_) + ADDON);
    if (qwait != 0) {
      eta = etaq;
      qprime = q;
      set_eta();
    } else {
      qwait = 2;
      compute_etaqm1();
      compute_etaqp1();
      choose_eta();
      set_eta();
    }
  }

  // cvBDFStab (cvodes.c:8158-8219): at order q >= 3 the squared scaled
  // derivative norms of orders q-1, q and q+1 of this step are pushed onto the
  // five-step history; when no order increase is planned and q + 5 steps have
  // passed since the last change of step size (nscon), the detector decides
  // whether the order is stability-limited, and if so the next step is taken
  // at order q-1 with that order's step ratio.  An order increase restarts the
  // count.
  __device__ __forceinline__ void bdf_stab() {
    if (q >= 3) {
      // the squares of sqm2 = (q-1)! |zn[q-1]|, sqm1 = q! |zn[q]| (WRMS
      // norms) and sq = (q+1)! acnrm / tq[5], formed from the sums of squares
      // without the square roots
      double fact = 1.0;
      for (int i = 2; i < q; ++i) fact *= i;
      const double fq = fact * q;
      const double e1 = active ? Z(q) * ewt : 0.0;
      const double e2 = active ? Z(q - 1) * ewt : 0.0;
      const double s1 = tm.sum(e1 * e1), s2 = tm.sum(e2 * e2);
      constexpr double RNX = 1.0 / NX;
      const double sq = fq * (q + 1) * stald_div(acnrm, fmax(S.tq[5], 1.0e-10));
      double ss[4][3];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int k = 0; k < 3; ++k) ss[i][k] = S.ssdat[i][k];
      team_sync();
      if (tm.lane == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) S.ssdat[i + 1][k] = ss[i][k];
        S.ssdat[0][0] = fact * fact * (s2 * RNX);
        S.ssdat[0][1] = fq * fq * (s1 * RNX);
        S.ssdat[0][2] = sq * sq;
      }
      team_sync();
    }
    if (qprime >= q) {
      if (q >= 3 && nscon >= q + 5) {
        double ss[5][3];
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
          for (int k = 0; k < 3; ++k) ss[i][k] = S.ssdat[i][k];
        if (stald_flag(ss, q) > 3) {
          qprime = q - 1;
          eta = fmin(etaqm1, etamax);
          hprime = h * eta;
          nor++;
        }
      }
    } else {
      nscon = 0;
    }
  }

  // ---------------- one step: cvStep ----------------
  __device__ __forceinline__ int step() {
    int ncf = 0, nef = 0, nefQ = 0, ncfS = 0, nefS = 0;
    if (nst > 0 && hprime != h) adjust_params();
    const double saved_t = tn;
    int nflag = FIRST_CALL;
    double dsm = 0.0;
    int attempt = 0;
    for (;; ++attempt) {
      if (attempt >= MAX_ATTEMPTS) return AMX_CONV_FAILURE;
      AMX_PF_BEGIN(t0);
      predict();
      AMX_PF_END(t0, PF_PREDICT);
      AMX_PF_BEGIN(t1);
      set_coeffs();
      AMX_PF_END(t1, PF_COEFF);
      AMX_PF_BEGIN(t2);
      nflag = nls(nflag);
      AMX_PF_END(t2, PF_NEWTON);
      int kflag = handle_nflag(nflag, saved_t, ncf);
      if (kflag == PREDICT_AGAIN) continue;
      if (kflag != DO_ERROR_TEST) return kflag;
      AMX_PF_BEGIN(t3);
      int eflag = error_test(nflag, saved_t, acnrm, nef);
      AMX_PF_END(t3, PF_ERRTEST);
      if (eflag == TRY_AGAIN) continue;
      if (eflag != AMX_SUCCESS) return eflag;
      dsm = dsm_test;
      if constexpr (QUAD) {
        ncf = nef = 0;
        AMX_PF_BEGIN(t4);
        quad_correct();
        AMX_PF_END(t4, PF_QUAD);
        if (quad_errcon()) {
          AMX_PF_BEGIN(t5);
          const double acnrmQ = wrmsQ(acorQ, ewtQ);
          eflag = error_test(nflag, saved_t, acnrmQ, nefQ);
          AMX_PF_END(t5, PF_ERRTEST);
          if (eflag == TRY_AGAIN) continue;
          if (eflag != AMX_SUCCESS) return eflag;
          dsm = fmax(dsm, dsm_test);
        }
      }
      if constexpr (SENS) {
        ncf = nef = 0;
        nefQ = 0;
        ftemp = pb.rhs(tn, y);
        nfe++;
        int sflag = pb.sens_nls(*this);
        kflag = handle_nflag(sflag, saved_t, ncfS);
        nflag = sflag;
        if (kflag == PREDICT_AGAIN) continue;
        if (kflag != DO_ERROR_TEST) return kflag;
        eflag = error_test(nflag, saved_t, acnrmS, nefS);
        if (eflag == TRY_AGAIN) continue;
        if (eflag != AMX_SUCCESS) return eflag;
        dsm = fmax(dsm, dsm_test);
      }
      break;
    }
    AMX_PF_BEGIN(t6);
    complete_step();
    AMX_PF_END(t6, PF_COMPLETE);
    record(1, dsm);
    AMX_PF_BEGIN(t7);
    prepare_next_step(dsm);
    if (sp.stldet) bdf_stab();
    AMX_PF_END(t7, PF_PREPARE);
    etamax = (nst <= SMALL_NST) ? ETA_MAX_ES : ETA_MAX_GS;
    dsm_last = dsm;
    return AMX_SUCCESS;
  }

  // ---------------- initial step size: cvHin ----------------
  __device__ __forceinline__ double upper_bound_h0(double tdist) {
    const double t1 = active ? (HUB_FACTOR * fabs(Z(0)) + 1.0 / ewt) : 1.0;
    double hub_inv = tm.max(active ? fabs(Z(1)) / t1 : 0.0);
    if constexpr (QUAD) {
      if (quad_errcon()) {
        double m = 0.0;
#pragma unroll
        for (int e = 0; e < EPQ; ++e)
          if (qslot(e) < nq)
            m = fmax(m, fabs(ZQ(1, e)) / (HUB_FACTOR * fabs(ZQ(0, e)) + 1.0 / ewtQ[e]));
        hub_inv = fmax(hub_inv, tm.max(m));
      }
    }
    if constexpr (SENS) hub_inv = fmax(hub_inv, pb.sens_hub_inv());
    double hub = HUB_FACTOR * tdist;
    if (hub * hub_inv > 1.0) hub = 1.0 / hub_inv;
    return hub;
  }

  __device__ __forceinline__ double ydd_norm(double hg) {
    const double yy = Z(0) + hg * Z(1);
    if constexpr (SENS) pb.sens_ydd_prepare(hg);
    const double fv = pb.rhs(tn + hg, yy);
    nfe++;
    const double tv = (fv - Z(1)) / hg;
    double r = wrms(tv, ewt);
    if constexpr (QUAD) {
      if (quad_errcon()) {
        double fq[EPQ];
        pb.quad_rhs(tn + hg, yy, fq);
        nfQe++;
        double tQ[EPQ];
#pragma unroll
        for (int e = 0; e < EPQ; ++e) tQ[e] = (fq[e] - ZQ(1, e)) / hg;
        r = fmax(r, wrmsQ(tQ, ewtQ));
      }
    }
    if constexpr (SENS) r = fmax(r, pb.sens_ydd_norm(tn + hg, yy, fv, hg));
    return r;
  }

  __device__ __forceinline__ int hin(double tout) {
    const double tdiff = tout - tn;
    if (tdiff == 0.0) return AMX_TOO_CLOSE;
    const double sign = (tdiff > 0.0) ? 1.0 : -1.0;
    const double tdist = fabs(tdiff);
    const double tround = UROUND * fmax(fabs(tn), fabs(tout));
    if (tdist < 2.0 * tround) return AMX_TOO_CLOSE;
    const double hlb = HLB_FACTOR * tround;
    const double hub = upper_bound_h0(tdist);
    double hg = sqrt(hlb * hub);
    if (hub < hlb) {
      h = sign * hg;
      return AMX_SUCCESS;
    }
    double hnew = hg, hs = hg;
    for (int count1 = 1; count1 <= 4; ++count1) {
      double yddnrm = 0.0;
      bool ok = false;
      for (int count2 = 1; count2 <= 4; ++count2) {
        yddnrm = ydd_norm(hg * sign);
        if (isfinite(yddnrm)) { ok = true; break; }
        hg *= 0.2;
      }
      if (!ok) {
        if (count1 <= 2) return AMX_RHSFUNC_FAIL;
        hnew = hs;
        break;
      }
      hs = hg;
      hnew = (yddnrm * hub * hub > 2.0) ? sqrt(2.0 / yddnrm) : sqrt(hg * hub);
      if (count1 == 4) break;
      const double hrat = hnew / hg;
      if (hrat > 0.5 && hrat < 2.0) break;
      if (count1 > 1 && hrat > 2.0) { hnew = hg; break; }
      hg = hnew;
    }
    double h0 = H_BIAS * hnew;
    if (h0 < hlb) h0 = hlb;
    if (h0 > hub) h0 = hub;
    h = sign * h0;
    return AMX_SUCCESS;
  }

  // ---------------- output: zn[0] or the Nordsieck interpolant ----------------
  __device__ __forceinline__ void out_zn0() {
    yout = Z(0);
#pragma unroll
    for (int e = 0; e < EPQ; ++e) yQout[e] = ZQ(0, e);
  }
  // y(t) = sum_j zn[j] s^j with s = (t - tn) / h (CVodeGetDky / CVodeGetQuadDky
  // with k = 0, cvodes.c:3742-3820, 3864-...): terms from j = q down to 0.
  __device__ __forceinline__ void out_dky(double t) {
    const double s = (t - tn) / h;
    double c[LMAX];
    c[0] = 1.0;
#pragma unroll
    for (int j = 1; j < LMAX; ++j) c[j] = c[j - 1] * s;
    double acc = 0.0;
    double accQ[EPQ];
#pragma unroll
    for (int e = 0; e < EPQ; ++e) accQ[e] = 0.0;
#pragma unroll
    for (int j = QMAX; j >= 0; --j) {
      if (j <= q) {
        acc = (j == q) ? c[j] * Z(j) : fma(c[j], Z(j), acc);
        if constexpr (QUAD)
#pragma unroll
          for (int e = 0; e < EPQ; ++e)
            accQ[e] = (j == q) ? c[j] * ZQ(j, e) : fma(c[j], ZQ(j, e), accQ[e]);
      }
    }
    yout = acc;
#pragma unroll
    for (int e = 0; e < EPQ; ++e) yQout[e] = accQ[e];
  }

  // ---------------- CVode(tout, itask) with tstop set ----------------
  // cvodes.c:2890-3620.  NORMAL mode (one_step false) returns at tstop with
  // zn[0], or, once a step reaches or passes tout, with the interpolant at
  // tout; ONE_STEP mode returns after one step.  The returned values are in
  // yout / yQout (at_tstop tells which case).  mxstep bounds the steps of
  // this call (CVodeSetMaxNumSteps).  pb.on_step() is invoked for the
  // initial point and after every accepted step (trajectory storage);
  // returning false stops with AMX_TRAJ_FULL.
  __device__ __forceinline__ int advance(double tout, double tstop_, int mxstep,
                                         bool one_step = false) {
    tstop = tstop_;
    tstopset = true;
    at_tstop = false;
    if (nst == 0) {
      set_ewt();
      if constexpr (QUAD) set_ewtQ();
      if constexpr (SENS) pb.sens_set_ewt();
      const double f0 = pb.rhs(tn, Z(0));
      nfe++;
      setZ(1, f0);
      if constexpr (QUAD) {
        double fq[EPQ];
        pb.quad_rhs(tn, Z(0), fq);
        nfQe++;
#pragma unroll
        for (int e = 0; e < EPQ; ++e) setZQ(1, e, fq[e]);
      }
      if constexpr (SENS) pb.sens_init_deriv(tn, Z(0), f0);
      if (any_nonfinite(f0)) return AMX_RHSFUNC_FAIL;
      if ((tstop - tn) * (tout - tn) <= 0.0) return AMX_ILL_INPUT;
      const double tout_hin = ((tout - tn) * (tout - tstop) > 0.0) ? tstop : tout;
      const int hf = hin(tout_hin);
      if (hf != AMX_SUCCESS) return hf;
      if ((tn + h - tstop) * h > 0.0) h = (tstop - tn) * (1.0 - 4.0 * UROUND);
      hscale = h;
      hprime = h;
      setZ(1, Z(1) * h);
      if constexpr (QUAD)
#pragma unroll
        for (int e = 0; e < EPQ; ++e) setZQ(1, e, ZQ(1, e) * h);
      if constexpr (SENS) pb.sens_scale_deriv(h);
      if (!pb.on_step(tn, Z(0), Z(1), hscale, 0)) return AMX_TRAJ_FULL;
    } else {
      const double troundoff = FUZZ_FACTOR * UROUND * (fabs(tn) + fabs(h));
      if (fabs(tn - tstop) <= troundoff) {
        if ((tout - tstop) * h >= 0.0 || fabs(tout - tstop) <= troundoff) {
          tstopset = false;
          at_tstop = true;
          out_zn0();
          return AMX_SUCCESS;
        }
      } else if ((tn + hprime - tstop) * h > 0.0) {
        hprime = (tstop - tn) * (1.0 - 4.0 * UROUND);
        eta = hprime / h;
      }
      if (!one_step && (tn - tout) * h >= 0.0) {
        out_dky(tout);
        return AMX_SUCCESS;
      }
    }
    for (int nstloc = 0;; ++nstloc) {
      AMX_PF_BEGIN(tl);
      if (nst > 0) {
        set_ewt();
        if constexpr (QUAD) if (quad_errcon()) set_ewtQ();
        if constexpr (SENS) pb.sens_set_ewt();
      }
      if (nstloc >= mxstep) return AMX_TOO_MUCH_WORK;
      if (sp.deadline != 0 && wall_clock64() > sp.deadline) return AMX_MAX_TIME_EXCEEDED;
      // CVode's "too much accuracy" test, tolsf = UROUND * WRMS(zn[0]) > 1.  With
      // the weights ewt = 1 / (rtol |zn0| + atol) just computed from zn[0]
      // (atol > 0) every weighted component is below 1 / rtol, so the test
      // can only fail when a relative tolerance is below UROUND; the norms are
      // computed only then, and otherwise a non-finite state is caught by a
      // ballot.
      if (sp.rtol < UROUND || (QUAD && quad_errcon() && sp.rtolQ < UROUND) ||
          (SENS && sp.rtolS < UROUND)) {
        double nrm = wrms(Z(0), ewt);
        if constexpr (QUAD) {
          if (quad_errcon()) {
            double z0[EPQ];
            znQ0(z0);
            nrm = fmax(nrm, wrmsQ(z0, ewtQ));
          }
        }
        if constexpr (SENS) nrm = fmax(nrm, pb.sens_norm_zn(0));
        if (!(nrm == nrm)) return AMX_ERR_FAILURE;
        tolsf = UROUND * nrm;
        if (tolsf > 1.0) return AMX_TOO_MUCH_ACC;
      } else if (any_nonfinite(Z(0))) {
        return AMX_ERR_FAILURE;
      }
      tolsf = 1.0;
      AMX_PF_END(tl, PF_LOOP);
      AMX_PF_BEGIN(ts);
      const int kflag = step();
      AMX_PF_END(ts, PF_STEP);
      if (kflag != AMX_SUCCESS) return kflag;
      const double troundoff = FUZZ_FACTOR * UROUND * (fabs(tn) + fabs(h));
      if (fabs(tn - tstop) <= troundoff) tn = tstop;
      if (!pb.on_step(tn, Z(0), Z(1), hscale, q)) return AMX_TRAJ_FULL;
      if constexpr (CK_SAVE) {
        // a checkpoint (CVodeF, cvodea.c:577-597): the state after this step,
        // and a fresh Newton matrix for the next one
        if (pb.ck_due >= 0) {
          save_state(pb.ck_rec + (size_t)pb.ck_due * CK_LEN, pb.jout);
          pb.ck_due = -1;
          force_setup = true;
        }
      }
      if (fabs(tn - tstop) <= troundoff) {
        if ((tout - tstop) * h >= 0.0 || fabs(tout - tstop) <= troundoff) {
          tstopset = false;
          at_tstop = true;
          out_zn0();
          return AMX_SUCCESS;
        }
      } else if ((tn + hprime - tstop) * h > 0.0) {
        hprime = (tstop - tn) * (1.0 - 4.0 * UROUND);
        eta = hprime / h;
      }
      if (!one_step && (tn - tout) * h >= 0.0) {
        out_dky(tout);
        return AMX_SUCCESS;
      }
      if (one_step) {
        out_zn0();
        return AMX_SUCCESS;
      }
    }
  }

  // ---------------- restart at a data point (AMICI's backward problem) ----------------
  // CVodeSolver::reset_state (src/solver_cvodes.cpp:471-505): order 1 and a
  // new zn[0] (and znQ[0], reinit_quad_b :635-644) at time t; every other
  // integrator variable keeps its value (step-size history, etamax, counters,
  // the linear-solver bookkeeping).
  __device__ __forceinline__ void reset_state(double t, double y0, const double (&q0)[EPQ]) {
    tn = t;
    q = 1;
    L_ = 2;
    qwait = L_;
    hu = 0.0;
    tolsf = 1.0;
    team_sync();
    setZ(0, y0);
    if constexpr (QUAD)
#pragma unroll
      for (int e = 0; e < EPQ; ++e) setZQ(0, e, q0[e]);
    team_sync();
  }
  // CVodeSolver::reInit_post_process (src/solver_cvodes.cpp:527-560): one
  // ONE_STEP call towards tnext with nst set to 0, so CVode's first-call
  // initialisation runs (f0, initial step from cvHin, and, since nst is 0 in
  // that step, a fresh Jacobian and Newton matrix with nstlp = nstlj = 0);
  // afterwards nst continues from its old value + 1.  A step that already
  // lands on tnext is an error (CV_TSTOP_RETURN is not CV_SUCCESS there).
  // With checkpoints the step also stops at the start of the current
  // checkpoint interval, tst >= tnext; landing there is no error.
  __device__ __forceinline__ int reinit_step(double tnext, int mxstep) {
    const int nst_tmp = nst;
    nst = 0;
    const int r = advance(tnext, tnext, mxstep, true);
    if (r != AMX_SUCCESS) return r;
    if (at_tstop) return AMX_ERROR;
    nst = nst_tmp + 1;
    return AMX_SUCCESS;
  }
  __device__ __forceinline__ int reinit_step(double tnext, double tst, int mxstep) {
    const int nst_tmp = nst;
    nst = 0;
    const int r = advance(tnext, tst, mxstep, true);
    if (r != AMX_SUCCESS) return r;
    if (at_tstop && tst == tnext) return AMX_ERROR;
    nst = nst_tmp + 1;
    return AMX_SUCCESS;
  }

  // ---------------- checkpoints (CVAckpntNew / CVAckpntGet, cvodea.c:1876-1990, 2381-2530) ----
  // The integrator state that the steps after this point depend on: the
  // Nordsieck array, the step-size history and method coefficients, the
  // stability-limit data, and every scalar carried from step to step; jout is
  // the caller's output index.  The Newton matrix is not part of it: the
  // caller forces a fresh one after the checkpoint, in the original pass and
  // in a recomputation alike, so both take the same steps.
  __device__ __forceinline__ void save_state(double* r, int jout) const {
    static_assert(!QUAD && !SENS, "checkpoints of the state problem only");
    if (active)
#pragma unroll
      for (int j = 0; j < LMAX; ++j) r[j * NX + tm.lane] = S.zn[j][tm.lane];
    if (tm.lane == 0) {
      double* c = r + LMAX * NX;
      int k = 0;
      for (int j = 0; j <= LMAX; ++j) c[k++] = S.tau[j];
      for (int j = 0; j < LMAX; ++j) c[k++] = S.l[j];
      for (int j = 0; j < 6; ++j) c[k++] = S.tq[j];
      for (int i = 0; i < 5; ++i)
        for (int m = 0; m < 3; ++m) c[k++] = S.ssdat[i][m];
      const double v[] = {tn, h, hprime, hscale, eta, etamax, etaq, etaqm1, etaqp1, gamma,
                          gammap, gamrat, rl1, crate, crateS, delp, acnrm, acnrmS, saved_tq5,
                          hu, tolsf, tstop, dsm_last, dsm_test,
                          (double)q, (double)qprime, (double)L_, (double)qwait, (double)nst,
                          (double)nstlp, (double)nstlj, (double)nscon, (double)convfail,
                          jcur ? 1.0 : 0.0, tstopset ? 1.0 : 0.0, at_tstop ? 1.0 : 0.0,
                          (double)jout};
      static_assert(7 + 6 + 6 + 15 + sizeof(v) / sizeof(double) <= CK_SCAL, "record size");
      for (unsigned m = 0; m < sizeof(v) / sizeof(double); ++m) c[k++] = v[m];
    }
  }
  // Restores a state written by save_state (counters start at 0); returns jout.
  __device__ __forceinline__ int load_state(const double* r) {
    team_sync();
    if (active)
#pragma unroll
      for (int j = 0; j < LMAX; ++j) S.zn[j][tm.lane] = r[j * NX + tm.lane];
    const double* c = r + LMAX * NX;
    if (tm.lane == 0) {
      int k = 0;
      for (int j = 0; j <= LMAX; ++j) S.tau[j] = c[k++];
      for (int j = 0; j < LMAX; ++j) S.l[j] = c[k++];
      for (int j = 0; j < 6; ++j) S.tq[j] = c[k++];
      for (int i = 0; i < 5; ++i)
        for (int m = 0; m < 3; ++m) S.ssdat[i][m] = c[k++];
    }
    const double* v = c + 7 + 6 + 6 + 15;
    tn = v[0]; h = v[1]; hprime = v[2]; hscale = v[3]; eta = v[4]; etamax = v[5];
    etaq = v[6]; etaqm1 = v[7]; etaqp1 = v[8]; gamma = v[9]; gammap = v[10]; gamrat = v[11];
    rl1 = v[12]; crate = v[13]; crateS = v[14]; delp = v[15]; acnrm = v[16]; acnrmS = v[17];
    saved_tq5 = v[18]; hu = v[19]; tolsf = v[20]; tstop = v[21]; dsm_last = v[22];
    dsm_test = v[23];
    q = (int)v[24]; qprime = (int)v[25]; L_ = (int)v[26]; qwait = (int)v[27];
    nst = (int)v[28]; nstlp = (int)v[29]; nstlj = (int)v[30]; nscon = (int)v[31];
    convfail = (int)v[32];
    jcur = v[33] != 0.0; tstopset = v[34] != 0.0; at_tstop = v[35] != 0.0;
    const int jout = (int)v[36];
    team_sync();
    nfe = nsetups = nje = nni = netf = ncfn = nfQe = nniS = nor = 0;
    y = Z(0);
    yout = y;
    ewt = 0.0; acor = 0.0; ftemp = 0.0;
    return jout;
  }
};

}  // namespace amx
