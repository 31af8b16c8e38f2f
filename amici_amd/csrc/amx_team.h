// Team primitives for gfx950: a "team" integrates one ODE instance.  Two
// kinds:
//  * T = 64: one whole 64-wide wavefront.  Teams of one block never wait
//    for each other, so no workgroup barrier is used.
//  * T = 128 or 256: the whole workgroup (T / 64 wavefronts, one team per
//    block), for models with more than 64 states.  Cross-wavefront
//    reductions and exchanges go through LDS with workgroup barriers.
// Teams narrower than a wavefront are not provided: their reductions would
// have to separate the teams sharing a wavefront, and every reduction here
// is built so that its result is one wave-uniform (scalar) value.
// All team-level control flow is uniform inside a team.  A translation unit
// holds one kind of multi-wavefront team: AMX_TEAM_LANES names its size.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifndef AMX_TEAM_LANES
#define AMX_TEAM_LANES 64
#endif

namespace amx {

// Workgroup barrier that orders LDS accesses only: the fences name the local
// address space, so outstanding global loads and stores stay in flight across
// it (a full __syncthreads() waits for them).  Used where the waves of a team
// exchange data through LDS; global data written by one lane and read by
// another needs __syncthreads().
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Orders a team's LDS traffic (the engine exchanges data between the lanes of
// a team only through LDS).
__device__ __forceinline__ void team_sync() {
#if AMX_TEAM_LANES > 64
  lds_barrier();
#else
  // A team never spans more than one wavefront, so wavefront-scope fences
  // order its LDS traffic: they keep the compiler from moving LDS accesses
  // across the point, and LDS requests of one wavefront are serviced in issue
  // order.  (Workgroup scope would also wait for outstanding global stores.)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#endif
}

// Value of v on wave lane `lane` (lane uniform).  The double is split into
// its two 32-bit halves with bit casts (no type-punned pointer access).
__device__ __forceinline__ double readlane_d(double v, int lane) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, lane);
  const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
  const unsigned long long r = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  return __builtin_bit_cast(double, r);
}

// 64-bit value through a 32-bit cross-lane primitive, one half at a time.
template <class F>
__device__ __forceinline__ double split_d(double v, F&& f) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = f((int)(unsigned)u);
  const int hi = f((int)(unsigned)(u >> 32));
  const unsigned long long r = ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo;
  return __builtin_bit_cast(double, r);
}

// Value of lane partner(lane) inside each row of 16 lanes, through a DPP move
// (VALU, no LDS round trip).  Controls used: quad_perm [1,0,3,2] (lane ^ 1),
// quad_perm [2,3,0,1] (lane ^ 2), row_half_mirror (7 - lane within 8),
// row_mirror (15 - lane within 16).
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  return split_d(v, [](int x) { return dpp_i<CTRL>(x); });
}

// Full-wavefront reductions.  Each row of 16 lanes reduces with DPP moves
// (xor 1, xor 2, half mirror, mirror: every lane of the row combines the same
// operand pairs, so the row result is the same on all its lanes) and the four
// row results are combined from lanes 0, 16, 32 and 48 with readlane in a
// fixed order, so the result is one scalar value for the whole wavefront
// (uniform by construction).
struct Wave64 {
  template <class Op>
  static __device__ __forceinline__ double row_reduce(double v, Op op) {
    v = op(v, dpp_d<DPP_XOR1>(v));
    v = op(v, dpp_d<DPP_XOR2>(v));
    v = op(v, dpp_d<DPP_HALF_MIRROR>(v));
    v = op(v, dpp_d<DPP_MIRROR>(v));
    return v;
  }
  static __device__ __forceinline__ double sum(double v) {
    v = row_reduce(v, [](double a, double b) { return a + b; });
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
  }
  static __device__ __forceinline__ double max(double v) {
    v = row_reduce(v, [](double a, double b) { return fmax(a, b); });
    return fmax(fmax(readlane_d(v, 0), readlane_d(v, 16)),
                fmax(readlane_d(v, 32), readlane_d(v, 48)));
  }
  static __device__ __forceinline__ int maxi(int v) {
    int u;
    u = dpp_i<DPP_XOR1>(v); v = u > v ? u : v;
    u = dpp_i<DPP_XOR2>(v); v = u > v ? u : v;
    u = dpp_i<DPP_HALF_MIRROR>(v); v = u > v ? u : v;
    u = dpp_i<DPP_MIRROR>(v); v = u > v ? u : v;
    int r = __builtin_amdgcn_readlane(v, 0);
    u = __builtin_amdgcn_readlane(v, 16); r = u > r ? u : r;
    u = __builtin_amdgcn_readlane(v, 32); r = u > r ? u : r;
    u = __builtin_amdgcn_readlane(v, 48); r = u > r ? u : r;
    return r;
  }
  // Maximum and the lowest wave lane that holds it (a maximum reduction and
  // a ballot of the lanes equal to it); v must not be NaN.
  static __device__ __forceinline__ double argmax_lane(double v, int& idx) {
    const double m = max(v);
    const unsigned long long bits = __builtin_amdgcn_ballot_w64(v == m);
    idx = (int)__builtin_ctzll(bits | (1ull << 63));
    return m;
  }
  // true when the predicate holds on any lane of the wave (uniform)
  static __device__ __forceinline__ bool any(bool pred) {
    return __builtin_amdgcn_ballot_w64(pred) != 0;
  }
  // (value, index) maximum; ties resolve to the smaller index.
  static __device__ __forceinline__ double argmax(double v, int i, int& idx) {
#define AMX_ARGMAX_STEP(CTRL)                                \
    {                                                        \
      const double u = dpp_d<CTRL>(v);                       \
      const int j = dpp_i<CTRL>(i);                          \
      if (u > v || (u == v && j < i)) { v = u; i = j; }      \
    }
    AMX_ARGMAX_STEP(DPP_XOR1)
    AMX_ARGMAX_STEP(DPP_XOR2)
    AMX_ARGMAX_STEP(DPP_HALF_MIRROR)
    AMX_ARGMAX_STEP(DPP_MIRROR)
#undef AMX_ARGMAX_STEP
    double bv = readlane_d(v, 0);
    int bi = __builtin_amdgcn_readlane(i, 0);
#pragma unroll
    for (int row = 16; row < 64; row += 16) {
      const double rv = readlane_d(v, row);
      const int ri = __builtin_amdgcn_readlane(i, row);
      if (rv > bv || (rv == bv && ri < bi)) { bv = rv; bi = ri; }
    }
    idx = bi;
    return bv;
  }
};

// LDS exchange area of a multi-wavefront team (one team per workgroup).
struct TeamXchg {
  double d[AMX_TEAM_LANES > 64 ? AMX_TEAM_LANES : 1];
  int i[AMX_TEAM_LANES > 64 ? AMX_TEAM_LANES : 1];
};
__device__ __forceinline__ TeamXchg& team_xchg() {
  __shared__ TeamXchg x;
  return x;
}

template <int T>
struct Team {
  static_assert(T == 64 || ((T == 128 || T == 256) && T == AMX_TEAM_LANES),
                "team size: one wavefront (64) or one workgroup of 128 / 256 lanes");
  static constexpr int W = T > 64 ? T / 64 : 1;  // wavefronts per team
  int lane;  // lane inside the team, 0..T-1

  __device__ __forceinline__ Team() {
    if constexpr (T > 64) lane = threadIdx.x;
    else lane = __lane_id();
  }

  // Multi-wavefront combine: every wavefront's (uniform) partial is written
  // by its first lane, the team waits, every lane folds the W partials in
  // wavefront order (the same order on every lane, so the result is team
  // uniform), and the team waits again before the area can be reused.
  template <class Op>
  __device__ __forceinline__ double combine(double wv, Op op) const {
    TeamXchg& x = team_xchg();
    if (__lane_id() == 0) x.d[lane >> 6] = wv;
    lds_barrier();
    double r = x.d[0];
#pragma unroll
    for (int w = 1; w < W; ++w) r = op(r, x.d[w]);
    lds_barrier();
    return r;
  }

  __device__ __forceinline__ double sum(double v) const {
    if constexpr (T > 64) {
      return combine(Wave64::sum(v), [](double a, double b) { return a + b; });
    } else {
      return Wave64::sum(v);
    }
  }
  __device__ __forceinline__ double max(double v) const {
    if constexpr (T > 64) {
      return combine(Wave64::max(v), [](double a, double b) { return fmax(a, b); });
    } else {
      return Wave64::max(v);
    }
  }
  // true when pred holds on any lane of the team (team-uniform)
  __device__ __forceinline__ bool any(bool pred) const {
    if constexpr (T > 64) return maxi(pred ? 1 : 0) != 0;
    else return Wave64::any(pred);
  }
  __device__ __forceinline__ int maxi(int v) const {
    if constexpr (T > 64) {
      const int wv = Wave64::maxi(v);
      TeamXchg& x = team_xchg();
      if (__lane_id() == 0) x.i[lane >> 6] = wv;
      lds_barrier();
      int r = x.i[0];
#pragma unroll
      for (int w = 1; w < W; ++w) r = x.i[w] > r ? x.i[w] : r;
      lds_barrier();
      return r;
    } else {
      return Wave64::maxi(v);
    }
  }
  // (value, index) maximum; ties resolve to the smaller index.
  __device__ __forceinline__ double argmax(double v, int& idx) const {
    if constexpr (T > 64) {
      int wi;
      const double wv = Wave64::argmax(v, lane, wi);
      TeamXchg& x = team_xchg();
      if (__lane_id() == 0) {
        x.d[lane >> 6] = wv;
        x.i[lane >> 6] = wi;
      }
      lds_barrier();
      double bv = x.d[0];
      int bi = x.i[0];
#pragma unroll
      for (int w = 1; w < W; ++w) {
        const double rv = x.d[w];
        const int ri = x.i[w];
        if (rv > bv || (rv == bv && ri < bi)) { bv = rv; bi = ri; }
      }
      lds_barrier();
      idx = bi;
      return bv;
    } else {
      return Wave64::argmax_lane(v, idx);
    }
  }
  // Value of v on team lane src (src uniform within the team).
  __device__ __forceinline__ double bcast(double v, int src) const {
    if constexpr (T > 64) {
      TeamXchg& x = team_xchg();
      if (lane == src) x.d[0] = v;
      lds_barrier();
      const double r = x.d[0];
      lds_barrier();
      return r;
    } else {
      return readlane_d(v, src);
    }
  }
  __device__ __forceinline__ int bcast_i(int v, int src) const {
    if constexpr (T > 64) {
      TeamXchg& x = team_xchg();
      if (lane == src) x.i[0] = v;
      lds_barrier();
      const int r = x.i[0];
      lds_barrier();
      return r;
    } else {
      return __builtin_amdgcn_readlane(v, src);
    }
  }
  // Value of v on team lane src (src may differ per lane, 0 <= src < T).
  __device__ __forceinline__ double gather(double v, int src) const {
    if constexpr (T > 64) {
      TeamXchg& x = team_xchg();
      x.d[lane] = v;
      lds_barrier();
      const double r = x.d[src];
      lds_barrier();
      return r;
    } else {
      return __shfl(v, src, T);
    }
  }
};

}  // namespace amx
