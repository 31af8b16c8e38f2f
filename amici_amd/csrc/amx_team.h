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

__device__ __forceinline__ double readlane_d(double v, int lane) {
  int2 p = *reinterpret_cast<int2*>(&v);
  p.x = __builtin_amdgcn_readlane(p.x, lane);
  p.y = __builtin_amdgcn_readlane(p.y, lane);
  return *reinterpret_cast<double*>(&p);
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

  // Butterfly reductions: every lane of the team ends with the bitwise-same
  // value because each level combines the same pair of operands.
  __device__ __forceinline__ double sum(double v) const {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, T);
    return v;
  }
  __device__ __forceinline__ double max(double v) const {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, T));
    return v;
  }
  __device__ __forceinline__ int maxi(int v) const {
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) {
      int u = __shfl_xor(v, o, T);
      v = u > v ? u : v;
    }
    return v;
  }
  // (value, index) maximum; ties resolve to the smaller index.
  __device__ __forceinline__ double argmax(double v, int& idx) const {
    int i = lane;
#pragma unroll
    for (int o = T / 2; o > 0; o >>= 1) {
      double u = __shfl_xor(v, o, T);
      int j = __shfl_xor(i, o, T);
      if (u > v || (u == v && j < i)) { v = u; i = j; }
    }
    idx = i;
    return v;
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
