// Team primitives for gfx950: a "team" is T consecutive lanes of one 64-wide
// wavefront (T a power of two, 1..64) that integrates one ODE instance.  All
// team-level control flow is uniform inside a team; different teams of the
// same wavefront may diverge, so no workgroup barrier is ever used here.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amx {

__device__ __forceinline__ void team_sync() {
  // Orders this wavefront's LDS traffic (s_waitcnt) and stops the compiler
  // from moving LDS accesses across the point; LDS requests of one wavefront
  // are serviced in issue order, so this is sufficient inside a wavefront.
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
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

// Exchange with lane (lane ^ O) inside each group of 32 lanes: ds_swizzle in
// bit-mask mode (and 0x1f, or 0, xor O), an immediate pattern.
template <int O>
__device__ __forceinline__ double swz_xor(double v) {
  return split_d(v, [](int x) { return __builtin_amdgcn_ds_swizzle(x, (O << 10) | 0x1F); });
}
template <int O>
__device__ __forceinline__ int swz_xor_i(int x) {
  return __builtin_amdgcn_ds_swizzle(x, (O << 10) | 0x1F);
}

template <int T>
struct Team {
  static_assert(T >= 1 && T <= 64 && (T & (T - 1)) == 0, "team size");
  int lane;  // lane inside the team, 0..T-1
  int base;  // wave lane of team lane 0

  __device__ __forceinline__ Team() {
    const int wl = __lane_id();
    lane = wl & (T - 1);
    base = wl - lane;
  }

  // Reductions.  For a full-wave team (T = 64) each 32-lane half reduces
  // with immediate-pattern swizzles and the two half results are combined
  // from lanes 0 and 32 with readlane, so the result is one scalar value for
  // the whole wave (uniform by construction).  Smaller teams use a butterfly
  // of lane shuffles; every lane of the team then ends with the bitwise-same
  // value because each level combines the same pair of operands.
  template <class Op>
  __device__ __forceinline__ double half_reduce(double v, Op op) const {
    v = op(v, swz_xor<16>(v));
    v = op(v, swz_xor<8>(v));
    v = op(v, swz_xor<4>(v));
    v = op(v, swz_xor<2>(v));
    v = op(v, swz_xor<1>(v));
    return v;
  }
  __device__ __forceinline__ double sum(double v) const {
    if constexpr (T == 64) {
      v = half_reduce(v, [](double a, double b) { return a + b; });
      return readlane_d(v, 0) + readlane_d(v, 32);
    } else {
#pragma unroll
      for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
      return v;
    }
  }
  __device__ __forceinline__ double max(double v) const {
    if constexpr (T == 64) {
      v = half_reduce(v, [](double a, double b) { return fmax(a, b); });
      return fmax(readlane_d(v, 0), readlane_d(v, 32));
    } else {
#pragma unroll
      for (int o = T / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, T));
      return v;
    }
  }
  __device__ __forceinline__ int maxi(int v) const {
    if constexpr (T == 64) {
      int u;
      u = swz_xor_i<16>(v); v = u > v ? u : v;
      u = swz_xor_i<8>(v); v = u > v ? u : v;
      u = swz_xor_i<4>(v); v = u > v ? u : v;
      u = swz_xor_i<2>(v); v = u > v ? u : v;
      u = swz_xor_i<1>(v); v = u > v ? u : v;
      const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 32);
      return a > b ? a : b;
    } else {
#pragma unroll
      for (int o = T / 2; o > 0; o >>= 1) {
        int u = __shfl_xor(v, o, T);
        v = u > v ? u : v;
      }
      return v;
    }
  }
  // (value, index) maximum; ties resolve to the smaller index.
  __device__ __forceinline__ double argmax(double v, int& idx) const {
    int i = lane;
    if constexpr (T == 64) {
#define AMX_ARGMAX_LEVEL(O)                                  \
      {                                                      \
        const double u = swz_xor<O>(v);                      \
        const int j = swz_xor_i<O>(i);                       \
        if (u > v || (u == v && j < i)) { v = u; i = j; }    \
      }
      AMX_ARGMAX_LEVEL(16)
      AMX_ARGMAX_LEVEL(8)
      AMX_ARGMAX_LEVEL(4)
      AMX_ARGMAX_LEVEL(2)
      AMX_ARGMAX_LEVEL(1)
#undef AMX_ARGMAX_LEVEL
      const double v0 = readlane_d(v, 0), v1 = readlane_d(v, 32);
      const int i0 = __builtin_amdgcn_readlane(i, 0), i1 = __builtin_amdgcn_readlane(i, 32);
      const bool hi = (v1 > v0) || (v1 == v0 && i1 < i0);
      idx = hi ? i1 : i0;
      return hi ? v1 : v0;
    } else {
#pragma unroll
      for (int o = T / 2; o > 0; o >>= 1) {
        double u = __shfl_xor(v, o, T);
        int j = __shfl_xor(i, o, T);
        if (u > v || (u == v && j < i)) { v = u; i = j; }
      }
      idx = i;
      return v;
    }
  }
  // Value of v on team lane src (src uniform within the team).
  __device__ __forceinline__ double bcast(double v, int src) const {
    if constexpr (T == 64) return readlane_d(v, src);
    else return __shfl(v, src, T);
  }
  __device__ __forceinline__ int bcast_i(int v, int src) const {
    if constexpr (T == 64) return __builtin_amdgcn_readlane(v, src);
    else return __shfl(v, src, T);
  }
  // Value of v on team lane src (src may differ per lane).
  __device__ __forceinline__ double gather(double v, int src) const {
    return __shfl(v, src, T);
  }
};

}  // namespace amx
