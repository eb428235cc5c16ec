// Leapfrog kernels and the bounding-box reduction.
//
// kick_drift fuses the first kick and the drift of leapfrog_step
// (integrator.cu:5-33) and, for Barnes-Hut, the per-workgroup bounding boxes
// of the new positions (forces_bh.cu:125-156), so one pass over pos/vel/acc
// does all three; the Morton kernel finishes the reduction
// (forces_bh.cu:158-188).  The closing kick (integrator.cu:35) is fused into the force
// kernels' epilogues.  The FP32 operations are the reference's:
//   v += (dt/2) a;  x += dt v   (as fmaf), mass in pos.w untouched.
#include "nb_device.hpp"
#include "nb_internal.hpp"

#include <cfloat>

namespace nb {
namespace {

// Block reduction of the per-thread boxes; thread 0 writes the block's
// (lo, hi) pair.  The Morton kernel reduces the pairs (kernel boundary =
// visibility), so no cross-workgroup synchronisation is needed here.
__device__ void bbox_block_partial(float lo[3], float hi[3], float4* partial) {
  constexpr int W = kBlock / 64;
  __shared__ float s_lo[W][3], s_hi[W][3];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    lo[a] = wave_min(lo[a]);
    hi[a] = wave_max(hi[a]);
  }
  if (lane == 0)
    for (int a = 0; a < 3; ++a) {
      s_lo[w][a] = lo[a];
      s_hi[w][a] = hi[a];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    float bl[3], bh[3];
    for (int a = 0; a < 3; ++a) {
      bl[a] = s_lo[0][a];
      bh[a] = s_hi[0][a];
      for (int k = 1; k < W; ++k) {
        bl[a] = fminf(bl[a], s_lo[k][a]);
        bh[a] = fmaxf(bh[a], s_hi[k][a]);
      }
    }
    partial[2 * blockIdx.x] = make_float4(bl[0], bl[1], bl[2], 0.f);
    partial[2 * blockIdx.x + 1] = make_float4(bh[0], bh[1], bh[2], 0.f);
  }
}

__global__ __launch_bounds__(kBlock) void kick_drift_bbox_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, const float4* __restrict__ acc, int n,
    float half_dt, float dt, float4* partial) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 v = vel[i], a = acc[i], p = pos[i];
    v.x = fmaf(half_dt, a.x, v.x);
    v.y = fmaf(half_dt, a.y, v.y);
    v.z = fmaf(half_dt, a.z, v.z);
    p.x = fmaf(dt, v.x, p.x);
    p.y = fmaf(dt, v.y, p.y);
    p.z = fmaf(dt, v.z, p.z);
    vel[i] = v;
    pos[i] = p;
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  bbox_block_partial(lo, hi, partial);
}

__global__ __launch_bounds__(kBlock) void kick_drift_kernel(float4* __restrict__ pos,
                                                            float4* __restrict__ vel,
                                                            const float4* __restrict__ acc,
                                                            int n, float half_dt, float dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = vel[i], a = acc[i], p = pos[i];
  v.x = fmaf(half_dt, a.x, v.x);
  v.y = fmaf(half_dt, a.y, v.y);
  v.z = fmaf(half_dt, a.z, v.z);
  p.x = fmaf(dt, v.x, p.x);
  p.y = fmaf(dt, v.y, p.y);
  p.z = fmaf(dt, v.z, p.z);
  vel[i] = v;
  pos[i] = p;
}

__global__ __launch_bounds__(kBlock) void bbox_kernel(const float4* __restrict__ pos, int n,
                                                      float4* partial) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    float4 p = pos[i];
    lo[0] = fminf(lo[0], p.x); lo[1] = fminf(lo[1], p.y); lo[2] = fminf(lo[2], p.z);
    hi[0] = fmaxf(hi[0], p.x); hi[1] = fmaxf(hi[1], p.y); hi[2] = fmaxf(hi[2], p.z);
  }
  bbox_block_partial(lo, hi, partial);
}

__global__ __launch_bounds__(kBlock) void kick_kernel(float4* __restrict__ vel,
                                                      const float4* __restrict__ acc, int n,
                                                      float half_dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = vel[i], a = acc[i];
  v.x = fmaf(half_dt, a.x, v.x);
  v.y = fmaf(half_dt, a.y, v.y);
  v.z = fmaf(half_dt, a.z, v.z);
  vel[i] = v;
}

__global__ void fill_u32_kernel(unsigned* p, unsigned v, size_t count) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < count; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

}  // namespace

void warm_integrate() {
  hipFuncAttributes a;
  NB_HIP(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(kick_drift_bbox_kernel)));
}

int bbox_parts(int n) {
  int g = (n + kBlock - 1) / kBlock;
  return g < kBboxBlocks ? (g > 0 ? g : 1) : kBboxBlocks;
}

void launch_kick_drift(hipStream_t s, float4* pos, float4* vel, const float4* acc, int n,
                       float half_dt, float dt, bool with_bbox, float4* partial) {
  if (n <= 0) return;
  if (with_bbox) {
    kick_drift_bbox_kernel<<<bbox_parts(n), kBlock, 0, s>>>(pos, vel, acc, n, half_dt, dt,
                                                            partial);
  } else {
    kick_drift_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(pos, vel, acc, n, half_dt,
                                                                   dt);
  }
  NB_LAUNCH_CHECK();
}

void launch_bbox(hipStream_t s, const float4* pos, int n, float4* partial) {
  if (n <= 0) return;
  bbox_kernel<<<bbox_parts(n), kBlock, 0, s>>>(pos, n, partial);
  NB_LAUNCH_CHECK();
}

void launch_kick(hipStream_t s, float4* vel, const float4* acc, int n, float half_dt) {
  if (n <= 0) return;
  kick_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(vel, acc, n, half_dt);
  NB_LAUNCH_CHECK();
}

void launch_fill_u32(hipStream_t s, unsigned* p, unsigned v, size_t count) {
  if (count == 0) return;
  size_t g = (count + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  fill_u32_kernel<<<(unsigned)g, kBlock, 0, s>>>(p, v, count);
  NB_LAUNCH_CHECK();
}

}  // namespace nb
