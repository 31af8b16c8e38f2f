// Component kernels for the GPU tests: each one exercises a single building
// block of the integrator (team reductions / broadcasts, row-per-lane LU or
// the blocked global-memory LU of models with more than 64 states,
// mass-action model evaluation) on inputs from the host, with no integration
// loop, so results can be compared element by element with numpy.
// Compiled together with a generated model's traits (AMX_MODEL_TRAITS).
#include AMX_MODEL_TRAITS

namespace {
using M = AmxModel;
constexpr int T = M::T;
constexpr int NX = M::NX;
constexpr int TEAMS = amx::Launch<M>::TEAMS;
constexpr int THREADS = amx::Launch<M>::THREADS;

// out per lane: sum, max, argmax value, argmax index, bcast from lane 1 % T,
// gather from lane (T-1-lane)
__global__ void __launch_bounds__(THREADS) k_team(const double* in, double* out, int nteams) {
  amx::Team<T> tm;
  const int team = (blockIdx.x * blockDim.x + threadIdx.x) / T;
  if (team >= nteams) return;
  const double v = in[team * T + tm.lane];
  int idx = 0;
  const double s = tm.sum(v);
  const double mx = tm.max(v);
  const double am = tm.argmax(v, idx);
  const double b = tm.bcast(v, 1 % T);
  const double g = tm.gather(v, T - 1 - tm.lane);
  double* o = out + (size_t)(team * T + tm.lane) * 6;
  o[0] = s; o[1] = mx; o[2] = am; o[3] = idx; o[4] = b; o[5] = g;
}

// LU factor of A (row-major input, row i on lane i) and solve A x = b; models
// with more than 64 states factorise a column-major copy in `work` (LUBig)
template <class MM, bool BIG = amx::big_model<MM>()>
struct LuCase {  // row-per-lane LU in registers and LDS
  static __device__ __forceinline__ bool run(const amx::Team<T>& tm, const double* A, double bi,
                                             double* xi, double*, amx::BdfLds<NX, 0>& S) {
    const int i = tm.lane;
    amx::LURows<NX, T> lu;
    double a[NX];
#pragma unroll
    for (int j = 0; j < NX; ++j) a[j] = (i < NX) ? A[(size_t)i * NX + j] : 0.0;
    const bool good = lu.factor(tm, a, S);
    *xi = lu.solve(bi, tm, S);
    return good;
  }
};
template <class MM>
struct LuCase<MM, true> {  // blocked LU in global memory
  static __device__ __forceinline__ bool run(const amx::Team<T>& tm, const double* A, double bi,
                                             double* xi, double* W, amx::BdfLds<NX, 0>& S) {
    const int i = tm.lane;
    amx::LUBig<MM::NX, MM::T> lu;
    if (i < NX)
      for (int j = 0; j < NX; ++j) W[(size_t)j * NX + i] = A[(size_t)i * NX + j];
    amx::team_sync();
    const bool good = lu.factor(tm, W, S);
    *xi = lu.solve(bi, tm, W, S);
    return good;
  }
};

__global__ void __launch_bounds__(THREADS) k_lu(const double* A, const double* b, double* x, int* ok,
                                            double* work, int nteams) {
  amx::Team<T> tm;
  const int team = (blockIdx.x * blockDim.x + threadIdx.x) / T;
  if (team >= nteams) return;
  __shared__ amx::BdfLds<NX, 0> lds[TEAMS];
  amx::BdfLds<NX, 0>& S = lds[threadIdx.x / T];
  const int i = tm.lane;
  const double bi = (i < NX) ? b[(size_t)team * NX + i] : 0.0;
  double xi = 0.0;
  const bool good = LuCase<M>::run(tm, A + (size_t)team * NX * NX, bi, &xi,
                                                     work + (size_t)team * NX * NX, S);
  if (i < NX) x[(size_t)team * NX + i] = good ? xi : NAN;
  if (i == 0) ok[team] = good ? 1 : 0;
}

// model pieces at (theta, x, v): xdot, dense J, u = S^T v, J^T v (adjoint form)
__global__ void __launch_bounds__(THREADS) k_model(const double* theta, const double* x,
                                               const double* v, double* xdot, double* J,
                                               double* u, double* jtv, int nteams) {
  __shared__ amx::Smem<M> smem[TEAMS];
  amx::Team<T> tm;
  const int team = (blockIdx.x * blockDim.x + threadIdx.x) / T;
  if (team >= nteams) return;
  amx::Smem<M>& s = smem[threadIdx.x / T];
  for (int r = tm.lane; r < M::NR; r += T) s.pk[r] = exp10(theta[(size_t)team * M::NP + M::KI[r]]);
  amx::team_sync();
  amx::MAEval<M> ev(tm, s);
  const double xi = (tm.lane < NX) ? x[(size_t)team * NX + tm.lane] : 0.0;
  const double f = ev.xdot(xi);
  if (tm.lane < NX) xdot[(size_t)team * NX + tm.lane] = f;
  ev.put_x(xi);
  ev.jacobian();
  if constexpr (amx::big_model<M>()) {
    for (int e = tm.lane; e < NX * NX; e += T) J[(size_t)team * NX * NX + e] = 0.0;
    amx::team_sync();
    for (int e = tm.lane; e < M::JNZ; e += T)
      J[(size_t)team * NX * NX + M::J_I[e] * NX + M::J_J[e]] = s.Jv[e];
  } else {
    for (int e = tm.lane; e < NX * NX; e += T)
      J[(size_t)team * NX * NX + e] = s.J[(e / NX) * amx::Smem<M>::JLD + (e % NX)];
  }
  amx::team_sync();
  // adjoint RHS through the flux structure, as BwdProblem::rhs
  if (tm.lane < NX) s.v[tm.lane] = v[(size_t)team * NX + tm.lane];
  amx::team_sync();
  ev.stoich_t();
  for (int r = tm.lane; r < M::NR; r += T) u[(size_t)team * M::NR + r] = s.u[r];
  double acc = 0.0;
  if (tm.lane < NX)
    for (int q = M::DX_PTR[tm.lane]; q < M::DX_PTR[tm.lane + 1]; ++q) {
      const int r = M::DX_R[q], o = M::DX_O[q];
      double d = M::DX_C[q] * s.pk[r];
      if (o >= 0) d *= s.x[o];
      acc = fma(d, s.u[r], acc);
    }
  if (tm.lane < NX) jtv[(size_t)team * NX + tm.lane] = acc;
}

int blocks_for(int nteams) { return (nteams + TEAMS - 1) / TEAMS; }
}  // namespace

extern "C" {
int amxt_dims(int* d) {
  d[0] = NX; d[1] = M::NR; d[2] = M::NP; d[3] = T;
  return 0;
}
int amxt_team(const double* in, double* out, int nteams, void* stream) {
  hipLaunchKernelGGL(k_team, dim3(blocks_for(nteams)), dim3(THREADS), 0, (hipStream_t)stream, in, out,
                     nteams);
  return hipGetLastError() == hipSuccess ? 0 : -30;
}
int amxt_lu(const double* A, const double* b, double* x, int* ok, double* work, int nteams,
            void* stream) {
  hipLaunchKernelGGL(k_lu, dim3(blocks_for(nteams)), dim3(THREADS), 0, (hipStream_t)stream, A, b, x,
                     ok, work, nteams);
  return hipGetLastError() == hipSuccess ? 0 : -30;
}
int amxt_model(const double* theta, const double* x, const double* v, double* xdot, double* J,
               double* u, double* jtv, int nteams, void* stream) {
  hipLaunchKernelGGL(k_model, dim3(blocks_for(nteams)), dim3(THREADS), 0, (hipStream_t)stream, theta,
                     x, v, xdot, J, u, jtv, nteams);
  return hipGetLastError() == hipSuccess ? 0 : -30;
}
}
