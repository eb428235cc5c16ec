// Tiled all-pairs gravity and the FP64 energy diagnostic.
//
// forces_naive (forces.cu:33-60): every target sums the softened pull of
// every source, sources streamed through LDS in float4 tiles of TILE sources
// (256, 512 or 1024: one barrier pair per tile; the plain FP32 sums are the
// same bits at every tile width, the compensated ones Kahan-add per tile).
// Differences in this design:
//   * each thread owns kTargets targets (1: measured fastest on gfx950 --
//     the loop is VALU-issue bound at ~13 VALU per interaction, and more
//     targets per thread only raise register pressure);
//   * for small n the source range is split over blockIdx.y so the grid still
//     covers all CUs; the partial sums are added in a fixed order by a second
//     kernel (deterministic, no atomics);
//   * with `compensated`, every tile's FP32 partial is Kahan-added into the
//     running total (pairwise tile sums);
//   * the closing kick of the leapfrog step can be fused into the epilogue.
// direct_f64: FP64 direct summation for chosen targets, the accuracy
// reference the force error is measured against.
// energy (integrator.cu:43-110): per-target potential summed in FP32 within a
// tile and FP64 across tiles, self term excluded exactly, KE and PE reduced in
// FP64 per block and then in a fixed order by one block (no double atomics).
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

constexpr int kTargets = 1;
constexpr int kSpan = kBlock * kTargets;   // targets per block

template <bool COMP, int TILE>
__global__ __launch_bounds__(kBlock) void allpairs_kernel(const float4* __restrict__ pos, int n,
                                                          int chunk, float eps2, float G,
                                                          float4* __restrict__ out,
                                                          float4* __restrict__ vel,
                                                          float half_dt, int direct) {
  static_assert(TILE % kBlock == 0, "a tile is whole workgroup loads");
  __shared__ float4 tile[TILE];
  float bx[kTargets], by[kTargets], bz[kTargets];
  float sx[kTargets], sy[kTargets], sz[kTargets];
  float cx[kTargets], cy[kTargets], cz[kTargets];
  const int i0 = blockIdx.x * kSpan + threadIdx.x;
#pragma unroll
  for (int q = 0; q < kTargets; ++q) {
    int i = i0 + q * kBlock;
    float4 b = pos[i < n ? i : n - 1];
    bx[q] = b.x; by[q] = b.y; bz[q] = b.z;
    sx[q] = sy[q] = sz[q] = 0.f;
    cx[q] = cy[q] = cz[q] = 0.f;
  }
  const int s0 = blockIdx.y * chunk;
  const int s1 = min(n, s0 + chunk);
  for (int t = s0; t < s1; t += TILE) {
#pragma unroll
    for (int q = 0; q < TILE / kBlock; ++q) {
      const int j = t + q * kBlock + threadIdx.x;
      // massless padding: an exact +0 in every sum
      tile[q * kBlock + threadIdx.x] = j < s1 ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    __syncthreads();
    float px[kTargets], py[kTargets], pz[kTargets];
#pragma unroll
    for (int q = 0; q < kTargets; ++q) {
      if (COMP) { px[q] = py[q] = pz[q] = 0.f; }
      else { px[q] = sx[q]; py[q] = sy[q]; pz[q] = sz[q]; }
    }
#pragma unroll 8
    for (int k = 0; k < TILE; ++k) {
      float4 src = tile[k];
#pragma unroll
      for (int q = 0; q < kTargets; ++q)
        accumulate_src(src, bx[q], by[q], bz[q], eps2, px[q], py[q], pz[q]);
    }
#pragma unroll
    for (int q = 0; q < kTargets; ++q) {
      if (COMP) {
        kahan_add(sx[q], cx[q], px[q]);
        kahan_add(sy[q], cy[q], py[q]);
        kahan_add(sz[q], cz[q], pz[q]);
      } else {
        sx[q] = px[q]; sy[q] = py[q]; sz[q] = pz[q];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < kTargets; ++q) {
    int i = i0 + q * kBlock;
    if (i >= n) continue;
    if (direct) {
      float4 a = make_float4(G * sx[q], G * sy[q], G * sz[q], 0.f);
      out[i] = a;
      if (vel) {
        float4 v = vel[i];
        v.x = fmaf(half_dt, a.x, v.x);
        v.y = fmaf(half_dt, a.y, v.y);
        v.z = fmaf(half_dt, a.z, v.z);
        vel[i] = v;
      }
    } else {
      out[(size_t)blockIdx.y * n + i] = make_float4(sx[q], sy[q], sz[q], 0.f);
    }
  }
}

template <bool COMP>
__global__ __launch_bounds__(kBlock) void split_reduce_kernel(const float4* __restrict__ partial,
                                                              int n, int split, float G,
                                                              float4* __restrict__ acc,
                                                              float4* __restrict__ vel,
                                                              float half_dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float sx = 0.f, sy = 0.f, sz = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
  for (int s = 0; s < split; ++s) {
    float4 p = partial[(size_t)s * n + i];
    if (COMP) {
      kahan_add(sx, cx, p.x);
      kahan_add(sy, cy, p.y);
      kahan_add(sz, cz, p.z);
    } else {
      sx += p.x; sy += p.y; sz += p.z;
    }
  }
  float4 a = make_float4(G * sx, G * sy, G * sz, 0.f);
  acc[i] = a;
  if (vel) {
    float4 v = vel[i];
    v.x = fmaf(half_dt, a.x, v.x);
    v.y = fmaf(half_dt, a.y, v.y);
    v.z = fmaf(half_dt, a.z, v.z);
    vel[i] = v;
  }
}

__global__ __launch_bounds__(kBlock) void energy_kernel(const float4* __restrict__ pos,
                                                        const float4* __restrict__ vel, int n,
                                                        int chunk, float eps2, float G,
                                                        double* __restrict__ partial) {
  __shared__ float4 tile[kBlock];
  __shared__ double s_red[2][kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  float4 bi = pos[i < n ? i : n - 1];
  double phi = 0.0;
  const int s0 = blockIdx.y * chunk;
  const int s1 = min(n, s0 + chunk);
  for (int t = s0; t < s1; t += kBlock) {
    int j = t + threadIdx.x;
    tile[threadIdx.x] = j < s1 ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    float p = 0.f;
#pragma unroll 8
    for (int k = 0; k < kBlock; ++k) {
      float4 src = tile[k];
      float dx = src.x - bi.x, dy = src.y - bi.y, dz = src.z - bi.z;
      // self and padding terms are skipped by a select, not weighted by 0
      // (0 * rsq(0) would be NaN with eps = 0)
      const float term = fmaf(src.w, rsq(dist2(dx, dy, dz) + eps2), p);
      p = (t + k == i || t + k >= s1) ? p : term;
    }
    phi += (double)p;
    __syncthreads();
  }
  double ke = 0.0, pe = 0.0;
  if (i < n) {
    double m = (double)bi.w;
    if (blockIdx.y == 0) {
      float4 v = vel[i];
      double vx = v.x, vy = v.y, vz = v.z;
      ke = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    }
    pe = -0.5 * (double)G * m * phi;
  }
  ke = wave_sum(ke);
  pe = wave_sum(pe);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s_red[0][w] = ke;
    s_red[1][w] = pe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) {
      a += s_red[0][k];
      b += s_red[1][k];
    }
    size_t slot = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
    partial[2 * slot] = a;
    partial[2 * slot + 1] = b;
  }
}

__global__ __launch_bounds__(kBlock) void energy_final_kernel(const double* __restrict__ partial,
                                                              int count, double* out) {
  __shared__ double s_red[2][kBlock / 64];
  double ke = 0.0, pe = 0.0;
  for (int k = threadIdx.x; k < count; k += kBlock) {
    ke += partial[2 * k];
    pe += partial[2 * k + 1];
  }
  ke = wave_sum(ke);
  pe = wave_sum(pe);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s_red[0][w] = ke;
    s_red[1][w] = pe;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) {
      a += s_red[0][k];
      b += s_red[1][k];
    }
    out[0] = a + b;
    out[1] = a;
    out[2] = b;
  }
}

// FP64 direct summation for a list of targets (the accuracy reference of
// every force module): a_i = sum_{j != i} m_j d / (r2 * sqrt(r2)),
// r2 = |d|^2 + eps2, all in FP64 from the FP32 state, sources [lo, hi)
// streamed through LDS; blockIdx.y takes one source chunk and writes its
// partial sums, added in chunk order by direct_f64_reduce_kernel.
__global__ __launch_bounds__(kBlock) void direct_f64_kernel(const float4* __restrict__ pos,
                                                            int lo, int hi, int chunk,
                                                            const int* __restrict__ targets,
                                                            int nt, double eps2,
                                                            double* __restrict__ partial) {
  __shared__ float4 tile[kBlock];
  const int k = blockIdx.x * kBlock + threadIdx.x;
  const int i = targets[k < nt ? k : nt - 1];
  const float4 b = pos[i];
  const double bx = b.x, by = b.y, bz = b.z;
  double ax = 0.0, ay = 0.0, az = 0.0;
  const int s0 = lo + blockIdx.y * chunk;
  const int s1 = min(hi, s0 + chunk);
  for (int t = s0; t < s1; t += kBlock) {
    const int j = t + threadIdx.x;
    tile[threadIdx.x] = j < s1 ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    const int cnt = min(kBlock, s1 - t);
    for (int q = 0; q < cnt; ++q) {
      const float4 src = tile[q];
      const double dx = (double)src.x - bx, dy = (double)src.y - by, dz = (double)src.z - bz;
      const double r2 = dx * dx + dy * dy + dz * dz + eps2;
      const double f = (double)src.w / (r2 * sqrt(r2));
      if (t + q != i) {
        ax += dx * f;
        ay += dy * f;
        az += dz * f;
      }
    }
    __syncthreads();
  }
  if (k < nt) {
    double* o = partial + ((size_t)blockIdx.y * nt + k) * 3;
    o[0] = ax;
    o[1] = ay;
    o[2] = az;
  }
}

// One workgroup per target: thread t sums chunks t, t + 256, ... in order,
// then a fixed-order tree over the threads (deterministic for a given split).
__global__ __launch_bounds__(kBlock) void direct_f64_reduce_kernel(
    const double* __restrict__ partial, int nt, int split, double G, double* __restrict__ out) {
  __shared__ double s_a[3][kBlock];
  const int k = blockIdx.x;
  double a[3] = {0.0, 0.0, 0.0};
  for (int y = threadIdx.x; y < split; y += kBlock)
    for (int c = 0; c < 3; ++c) a[c] += partial[((size_t)y * nt + k) * 3 + c];
  for (int c = 0; c < 3; ++c) s_a[c][threadIdx.x] = a[c];
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int c = 0; c < 3; ++c) s_a[c][threadIdx.x] += s_a[c][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[(size_t)k * 3 + threadIdx.x] = G * s_a[threadIdx.x][0];
}

int split_for(int blocks_x, int n) {
  // aim for >= 2048 workgroups (8 per CU) but keep >= 4 tiles per chunk
  int max_split = n / (4 * kBlock);
  int split = (2048 + blocks_x - 1) / blocks_x;
  if (split > max_split) split = max_split;
  if (split > 64) split = 64;
  return split < 1 ? 1 : split;
}

int chunk_for(int n, int split, int tile = kBlock) {
  int c = (n + split - 1) / split;
  return ((c + tile - 1) / tile) * tile;
}

}  // namespace

void warm_allpairs() {
  hipFuncAttributes a;
  NB_HIP(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(energy_kernel)));
}

int allpairs_split(int n) { return split_for((n + kSpan - 1) / kSpan, n); }

void launch_allpairs(hipStream_t s, const float4* pos, int n, float eps2, float G, float4* acc,
                     float4* partial, int split, bool compensated, float4* vel_kick,
                     float half_dt, int tile) {
  if (n <= 0) return;
  if (tile == 0) tile = kBlock;
  if (tile != 256 && tile != 512 && tile != 1024)
    throw Error(kErrInvalid, "all-pairs tile must be 256, 512 or 1024 sources");
  if (split < 1) throw Error(kErrInvalid, "all-pairs split must be >= 1");
  int bx = (n + kSpan - 1) / kSpan;
  int chunk = chunk_for(n, split, tile);
  split = (n + chunk - 1) / chunk;
  dim3 grid(bx, split);
  bool direct = split == 1;
  float4* out = direct ? acc : partial;
  float4* kick = direct ? vel_kick : nullptr;
#define NB_AP(C, T) \
  allpairs_kernel<C, T><<<grid, kBlock, 0, s>>>(pos, n, chunk, eps2, G, out, kick, half_dt, direct)
  if (compensated) {
    if (tile == 256) NB_AP(true, 256);
    else if (tile == 512) NB_AP(true, 512);
    else NB_AP(true, 1024);
  } else {
    if (tile == 256) NB_AP(false, 256);
    else if (tile == 512) NB_AP(false, 512);
    else NB_AP(false, 1024);
  }
#undef NB_AP
  NB_LAUNCH_CHECK();
  if (!direct) {
    int g = (n + kBlock - 1) / kBlock;
    if (compensated)
      split_reduce_kernel<true><<<g, kBlock, 0, s>>>(partial, n, split, G, acc, vel_kick, half_dt);
    else
      split_reduce_kernel<false><<<g, kBlock, 0, s>>>(partial, n, split, G, acc, vel_kick, half_dt);
    NB_LAUNCH_CHECK();
  }
}

// Source chunks of the FP64 direct sum: enough workgroups to fill the chip
// (>= 2048) even for a few hundred targets, chunks of >= 4 LDS tiles.
static int direct_split_for(int bx, int nsrc) {
  int max_split = nsrc / (4 * kBlock);
  int split = (2048 + bx - 1) / bx;
  if (split > max_split) split = max_split;
  if (split > 4096) split = 4096;
  return split < 1 ? 1 : split;
}

int direct_f64_split(int nt, int nsrc) {
  const int bx = (nt + kBlock - 1) / kBlock;
  const int split = direct_split_for(bx, nsrc);
  const int chunk = chunk_for(nsrc, split);
  return (nsrc + chunk - 1) / chunk;
}

void launch_direct_f64(hipStream_t s, const float4* pos, int lo, int hi, const int* targets,
                       int nt, double eps2, double G, double* partial, double* out) {
  if (nt <= 0) return;
  const int nsrc = hi - lo;
  const int bx = (nt + kBlock - 1) / kBlock;
  const int split = direct_split_for(bx, nsrc);
  const int chunk = chunk_for(nsrc, split);
  const int ny = (nsrc + chunk - 1) / chunk;
  direct_f64_kernel<<<dim3(bx, ny), kBlock, 0, s>>>(pos, lo, hi, chunk, targets, nt, eps2,
                                                    partial);
  NB_LAUNCH_CHECK();
  direct_f64_reduce_kernel<<<nt, kBlock, 0, s>>>(partial, nt, ny, G, out);
  NB_LAUNCH_CHECK();
}

int energy_partials(int n) {
  int bx = (n + kBlock - 1) / kBlock;
  int split = split_for(bx, n);
  int chunk = chunk_for(n, split);
  split = (n + chunk - 1) / chunk;
  return bx * split;
}

void launch_energy(hipStream_t s, const float4* pos, const float4* vel, int n, float eps2,
                   float G, double* partial, int max_partial, double* out) {
  if (n <= 0) return;
  int bx = (n + kBlock - 1) / kBlock;
  int split = split_for(bx, n);
  int chunk = chunk_for(n, split);
  split = (n + chunk - 1) / chunk;
  if (bx * split > max_partial) throw Error(kErrState, "energy partial buffer too small");
  energy_kernel<<<dim3(bx, split), kBlock, 0, s>>>(pos, vel, n, chunk, eps2, G, partial);
  NB_LAUNCH_CHECK();
  energy_final_kernel<<<1, kBlock, 0, s>>>(partial, bx * split, out);
  NB_LAUNCH_CHECK();
}

}  // namespace nb
