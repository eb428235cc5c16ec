// Device helpers shared by the kernels: the canonical FP32 force law and
// wave64 reductions.  Everything is compiled with -ffp-contract=off, so an FMA
// appears only where fmaf is written; the CPU oracle uses the same order.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nb {

__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// Morton quantisation (forces_bh.cu:192-220), FP64 from the FP64 origin lo
// and the cubic scale inv = 1 / longest box edge:
//   q = clamp(trunc(((p - lo) * inv) * 2^21), 0, 2^21 - 1);
// spread21 interleaves q's 21 bits with two zero bits each.
__device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// spread10 interleaves v's low 10 bits with two zero bits each (32-bit ops).
// For a 21-bit q, spread10(q >> 11) is spread21(q) >> 33 on every third
// bit: the 30-bit key is the top 30 bits of the 63-bit key.
__device__ __forceinline__ uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | v << 16) & 0x030000ffu;
  v = (v | v << 8) & 0x0300f00fu;
  v = (v | v << 4) & 0x030c30c3u;
  v = (v | v << 2) & 0x09249249u;
  return v;
}

__device__ __forceinline__ uint32_t quant21(float p, double lo, double inv) {
  double v = ((double)p - lo) * inv * 2097152.0;
  v = (v > 0.0) ? v : 0.0;
  v = (v > 2097151.0) ? 2097151.0 : v;
  return (uint32_t)v;
}

// |d|^2 without softening, in the canonical order shared with the oracle.
__device__ __forceinline__ float dist2(float dx, float dy, float dz) {
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// Softened point-mass acceleration of one source (forces.cu:17-31), given the
// separation d = src - target and its d2 = dist2(d).
__device__ __forceinline__ void accumulate(float dx, float dy, float dz, float d2, float m,
                                           float eps2, float& ax, float& ay, float& az) {
  float inv = rsq(d2 + eps2);
  float f = m * inv * inv * inv;
  ax = fmaf(dx, f, ax);
  ay = fmaf(dy, f, ay);
  az = fmaf(dz, f, az);
}

// Same law for a source with no precomputed d2: the softening is folded
// into the first FMA of |d|^2 (one instruction fewer per interaction).
__device__ __forceinline__ void accumulate_src(float4 src, float bx, float by, float bz,
                                               float eps2, float& ax, float& ay, float& az) {
  const float dx = src.x - bx, dy = src.y - by, dz = src.z - bz;
  const float r2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2)));
  const float inv = rsq(r2);
  const float f = src.w * inv * inv * inv;
  ax = fmaf(dx, f, ax);
  ay = fmaf(dy, f, ay);
  az = fmaf(dz, f, az);
}

// Kahan-compensated add of `x` into (sum, c).
__device__ __forceinline__ void kahan_add(float& sum, float& c, float x) {
  float y = x - c;
  float t = sum + y;
  c = (t - sum) - y;
  sum = t;
}

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Number of set bits of the 64-bit mask m below this lane (two VALU).
__device__ __forceinline__ int mbcnt64(unsigned long long m) {
  return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also makes this
// wave's global stores visible to the workgroup, which waits for every store
// in flight (a round trip to L2) at each barrier; a loop that communicates
// through LDS and only streams results out to global memory does not need it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ __forceinline__ unsigned wave_incl_scan(unsigned v) {
  int lane = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// Agent-scope atomics (visible across all XCDs of the device).
__device__ __forceinline__ unsigned atomic_add_agent(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
}
// An LDS (address space 3) object.  Pointers into __shared__ storage that
// travel in structs or through selects can lose their address space, and
// accesses through a generic pointer are flat instructions, which wait on
// the global-memory counter as well; lds<T>* keeps them ds_* instructions.
// A __shared__ array converts with a C-style cast: (lds<T>*)array.  The host
// pass of a HIP compile only type-checks device code, so there it is T.
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using lds = __attribute__((address_space(3))) T;
#else
template <typename T>
using lds = T;
#endif

__device__ __forceinline__ unsigned atomic_add_wg(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ unsigned atomic_add_relaxed(unsigned* p, unsigned v) {
  return __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long load_agent(unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_agent_u32(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned load_agent_u32(unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace nb
