// Batched ensemble of independent systems stored back to back (system-major)
// in one engine: one set of launches per leapfrog step for all systems.
//
//   ens_kick_drift_bbox  fused first kick + drift and per-system bounding
//                        boxes (orderable-float atomics, one set per wave of
//                        256 consecutive particles per system it touches)
//   ens_morton_hist      per-system FP64 quantisation; key = system << 30 |
//                        30-bit Morton key (u64), digit histograms for the
//                        5 sort passes over bits 0..39.  Sorting by this key
//                        keeps every system contiguous, in its own Morton
//                        order, ties in original (= local) index order.
//   ens_karras           Karras build over the whole forest with the tie-break
//                        on the index *within* the system, so each system's
//                        subtree is exactly the tree a standalone build of
//                        that system gives; records each system's subtree
//                        root and resets the per-step bookkeeping.
//   ens_energy           per-system FP64 energy, segmented over target tiles.
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

__device__ __forceinline__ unsigned ord_of(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of(unsigned o) {
  return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

// system of sorted/original particle i: last s with off[s] <= i
__device__ __forceinline__ int sys_of(const int* off, int nsys, int i) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

constexpr int kEnsPer = 4;  // particles per thread in the bbox pass

__global__ __launch_bounds__(kBlock) void ens_kick_drift_bbox_kernel(
    float4* __restrict__ pos, float4* __restrict__ vel, const float4* __restrict__ acc, int n,
    float half_dt, float dt, int do_kick, const int* __restrict__ off, int nsys,
    unsigned* __restrict__ box) {
  const int lane = threadIdx.x & 63;
  const int wave_first = (blockIdx.x * blockDim.x + (threadIdx.x & ~63)) * kEnsPer;
  if (wave_first >= n) return;
  // the wave's 64 * kEnsPer consecutive particles, lane-fastest (coalesced)
  const int i0 = wave_first + lane;
  float4 p[kEnsPer];
#pragma unroll
  for (int k = 0; k < kEnsPer; ++k) {
    const int i = i0 + 64 * k;
    if (i < n) {
      p[k] = pos[i];
      if (do_kick) {
        float4 v = vel[i];
        const float4 a = acc[i];
        v.x = fmaf(half_dt, a.x, v.x);
        v.y = fmaf(half_dt, a.y, v.y);
        v.z = fmaf(half_dt, a.z, v.z);
        p[k].x = fmaf(dt, v.x, p[k].x);
        p[k].y = fmaf(dt, v.y, p[k].y);
        p[k].z = fmaf(dt, v.z, p[k].z);
        vel[i] = v;
        pos[i] = p[k];
      }
    }
  }
  const int wave_last = min(n - 1, wave_first + 64 * kEnsPer - 1);
  const int s_first = sys_of(off, nsys, wave_first);
  const int s_last = sys_of(off, nsys, wave_last);
  for (int s = s_first; s <= s_last; ++s) {
    const int a = off[s], b = off[s + 1];
    float lo[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
    float hi[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
#pragma unroll
    for (int k = 0; k < kEnsPer; ++k) {
      const int i = i0 + 64 * k;
      if (i < n && i >= a && i < b) {
        lo[0] = fminf(lo[0], p[k].x); lo[1] = fminf(lo[1], p[k].y); lo[2] = fminf(lo[2], p[k].z);
        hi[0] = fmaxf(hi[0], p[k].x); hi[1] = fmaxf(hi[1], p[k].y); hi[2] = fmaxf(hi[2], p[k].z);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      lo[c] = wave_min(lo[c]);
      hi[c] = wave_max(hi[c]);
    }
    if (lane == 0) {
      unsigned* bs = box + 8 * s;
      for (int c = 0; c < 3; ++c) {
        __hip_atomic_fetch_min(bs + c, ord_of(lo[c]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_max(bs + 3 + c, ord_of(hi[c]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
      }
    }
  }
}

constexpr int kEnsMaxSys = 1024;
// (system, key) words are below 2^40: five passes of 8-bit digits
constexpr int kEnsBits = kSortBits;
constexpr int kEnsPasses = 5;
constexpr int kEnsWords = kEnsPasses << kEnsBits;
static_assert(kEnsWords <= kHistWords && kEnsPasses * kEnsBits >= 40, "ensemble digit plan");

__global__ __launch_bounds__(kBlock) void ens_morton_hist_kernel(
    const float4* __restrict__ pos, int n, const int* __restrict__ off, int nsys,
    const unsigned* __restrict__ box, uint64_t* __restrict__ keys, SortState* ss) {
  __shared__ int s_off[kEnsMaxSys + 1];
  __shared__ double s_q[kEnsMaxSys][4];
  __shared__ unsigned h[kEnsWords];
  for (int i = threadIdx.x; i < kEnsWords; i += blockDim.x) h[i] = 0;
  for (int s = threadIdx.x; s <= nsys; s += blockDim.x) s_off[s] = off[s];
  for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
    // FP64 origin and cubic scale of system s (lbvh_check.py:45-48)
    double ext = 0.0;
    for (int c = 0; c < 3; ++c) {
      const float l = float_of(box[8 * s + c]), u = float_of(box[8 * s + 3 + c]);
      s_q[s][c] = (double)l;
      const double e = (double)u - (double)l;
      if (e > ext) ext = e;
    }
    s_q[s][3] = ext > 0.0 ? 1.0 / ext : 0.0;
  }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int s = sys_of(s_off, nsys, i);
    const float4 p = pos[i];
    const double inv = s_q[s][3];
    const uint64_t k63 = (spread21(quant21(p.x, s_q[s][0], inv)) << 2) |
                         (spread21(quant21(p.y, s_q[s][1], inv)) << 1) |
                         spread21(quant21(p.z, s_q[s][2], inv));
    const uint64_t k = ((uint64_t)s << 30) | (k63 >> 33);
    keys[i] = k;
#pragma unroll
    for (int q = 0; q < kEnsPasses; ++q)
      atomicAdd(&h[(q << kEnsBits) + ((unsigned)(k >> (kEnsBits * q)) & ((1u << kEnsBits) - 1u))],
                1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kEnsWords; i += blockDim.x) {
    const unsigned c = h[i];
    if (c) atomic_add_relaxed(&ss->hist[i], c);
  }
}

// forces_bh.cu:234-240 on (system, key) words with the tie-break on the
// index within the system (equal keys are always in the same system).
__device__ __forceinline__ int ens_delta(const uint64_t* __restrict__ keys, const int* off, int n,
                                         int i, uint64_t ki, int j) {
  if (j < 0 || j >= n) return -1;
  const uint64_t x = ki ^ keys[j];
  if (x) return __clzll((long long)x);
  const int base = off[(int)(ki >> 30)];
  return 64 + __clz((i - base) ^ (j - base));
}

__global__ __launch_bounds__(kBlock) void ens_karras_kernel(
    const uint64_t* __restrict__ keys, int n, const int* __restrict__ off, int nsys,
    Node* __restrict__ nodes, int2* __restrict__ up, unsigned* __restrict__ visit,
    int* __restrict__ roots, SortState* ss, unsigned* __restrict__ box) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.x == 0) {
    // this step's sort is done and its boxes consumed: reset for the next
    for (int k = threadIdx.x; k < kHistWords; k += blockDim.x) ss->hist[k] = 0;
    if (threadIdx.x < kMaxPasses) ss->ticket[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
      ss->epoch = ss->epoch + 1;
      ss->queue = 0;
    }
    for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
      for (int c = 0; c < 3; ++c) {
        box[8 * s + c] = 0xffffffffu;
        box[8 * s + 3 + c] = 0u;
      }
    }
  }
  if (i >= n - 1) return;
  const uint64_t ki = keys[i];
  const int d = (ens_delta(keys, off, n, i, ki, i + 1) - ens_delta(keys, off, n, i, ki, i - 1)) >= 0
                    ? 1
                    : -1;
  const int dmin = ens_delta(keys, off, n, i, ki, i - d);
  int lmax = 2;
  while (ens_delta(keys, off, n, i, ki, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t > 0; t >>= 1)
    if (ens_delta(keys, off, n, i, ki, i + (l + t) * d) > dmin) l += t;
  const int j = i + l * d;
  const int first = min(i, j), last = max(i, j);
  const uint64_t kf = keys[first];
  const int common = ens_delta(keys, off, n, first, kf, last);
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    if (split + step < last && ens_delta(keys, off, n, first, kf, split + step) > common)
      split += step;
  } while (step > 1);
  const int left = (split == first) ? (n - 1 + split) : split;
  const int right = (split + 1 == last) ? (n - 1 + split + 1) : split + 1;
  *reinterpret_cast<int2*>(&nodes[i].left) = make_int2(left, right);
  nodes[i].other = j;
  // A node whose range holds more than one system joins whole systems (its
  // split is a system boundary: the system bits differ there and nowhere
  // inside a system).  The walks start at the system roots and never read
  // such a node, so it gets no COM: its children's climbs end there as at a
  // root, and its pass word (3) keeps it out of every COM pass.
  const bool joins = (uint32_t)(kf >> 30) != (uint32_t)(keys[last] >> 30);
  if (joins) {
    up[left] = up_none();
    up[right] = up_none();
    visit[i] = 3u << 30;
  } else {
    const unsigned vw = com_visit_word(first, last);
    up[left] = make_up(i, vw, right);
    up[right] = make_up(i, vw, left);
    visit[i] = vw;
  }
  if (i == 0) up[0] = up_none();
  // the node covering exactly one system is that system's subtree root
  const int s = (int)(ki >> 30);
  if (first == off[s] && last == off[s + 1] - 1) roots[s] = i;
}

// Per-system energy: block b is target tile (system, first target) of up to
// 256 targets; sources are the system's particles.  FP32 pair potential per
// tile, FP64 across tiles, self term excluded exactly (integrator.cu:43-91).
__global__ __launch_bounds__(kBlock) void ens_energy_kernel(const float4* __restrict__ pos,
                                                            const float4* __restrict__ vel,
                                                            const int* __restrict__ off,
                                                            const int2* __restrict__ tiles,
                                                            float eps2, float G,
                                                            double* __restrict__ partial) {
  __shared__ float4 tile[kBlock];
  __shared__ double s_red[2][kBlock / 64];
  const int2 tl = tiles[blockIdx.x];
  const int s0 = off[tl.x], s1 = off[tl.x + 1];
  const int i = tl.y + threadIdx.x;
  const bool live = i < s1;
  const float4 bi = pos[live ? i : s1 - 1];
  double phi = 0.0;
  for (int t = s0; t < s1; t += kBlock) {
    const int j = t + threadIdx.x;
    tile[threadIdx.x] = j < s1 ? pos[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    __syncthreads();
    float p = 0.f;
#pragma unroll 8
    for (int k = 0; k < kBlock; ++k) {
      const float4 src = tile[k];
      const float dx = src.x - bi.x, dy = src.y - bi.y, dz = src.z - bi.z;
      const float term = fmaf(src.w, rsq(dist2(dx, dy, dz) + eps2), p);
      p = (t + k == i || t + k >= s1) ? p : term;  // self / padding skipped
    }
    phi += (double)p;
    __syncthreads();
  }
  double ke = 0.0, pe = 0.0;
  if (live) {
    const double m = (double)bi.w;
    const float4 v = vel[i];
    const double vx = v.x, vy = v.y, vz = v.z;
    ke = 0.5 * m * (vx * vx + vy * vy + vz * vz);
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
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = b;
  }
}

// One thread per system sums that system's tile partials in order.
__global__ void ens_energy_final_kernel(const double* __restrict__ partial,
                                        const int* __restrict__ tile_off, int nsys,
                                        double* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  double ke = 0.0, pe = 0.0;
  for (int b = tile_off[s]; b < tile_off[s + 1]; ++b) {
    ke += partial[2 * b];
    pe += partial[2 * b + 1];
  }
  out[3 * s] = ke + pe;
  out[3 * s + 1] = ke;
  out[3 * s + 2] = pe;
}

__global__ void ens_box_init_kernel(unsigned* box, int nsys) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  for (int c = 0; c < 3; ++c) {
    box[8 * s + c] = 0xffffffffu;
    box[8 * s + 3 + c] = 0u;
  }
}

}  // namespace

void warm_ensemble() {
  hipFuncAttributes a;
  NB_HIP(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(ens_box_init_kernel)));
}

int ens_passes() { return kEnsPasses; }
int ens_max_systems() { return kEnsMaxSys; }

void launch_ens_box_init(hipStream_t s, unsigned* box, int nsys) {
  ens_box_init_kernel<<<(nsys + kBlock - 1) / kBlock, kBlock, 0, s>>>(box, nsys);
  NB_LAUNCH_CHECK();
}

void launch_ens_kick_drift_bbox(hipStream_t s, float4* pos, float4* vel, const float4* acc,
                                int n, float half_dt, float dt, bool kick, const int* off,
                                int nsys, unsigned* box) {
  if (n <= 0) return;
  const int per_block = kBlock * kEnsPer;
  ens_kick_drift_bbox_kernel<<<(n + per_block - 1) / per_block, kBlock, 0, s>>>(
      pos, vel, acc, n, half_dt, dt, kick ? 1 : 0, off, nsys, box);
  NB_LAUNCH_CHECK();
}

void launch_ens_morton_hist(hipStream_t s, const float4* pos, int n, const int* off, int nsys,
                            const unsigned* box, uint64_t* keys, SortState* ss) {
  if (n <= 0) return;
  int g = (n + kBlock * 8 - 1) / (kBlock * 8);
  if (g > 2048) g = 2048;
  ens_morton_hist_kernel<<<g, kBlock, 0, s>>>(pos, n, off, nsys, box, keys, ss);
  NB_LAUNCH_CHECK();
}

void launch_ens_karras(hipStream_t s, const uint64_t* keys, int n, const int* off, int nsys,
                       Node* nodes, int2* up, unsigned* visit, int* roots, SortState* ss,
                       unsigned* box) {
  if (n < 2) return;
  const int g = (n - 1 + kBlock - 1) / kBlock;
  ens_karras_kernel<<<g, kBlock, 0, s>>>(keys, n, off, nsys, nodes, up, visit, roots, ss, box);
  NB_LAUNCH_CHECK();
}

void launch_ens_energy(hipStream_t s, const float4* pos, const float4* vel, const int* off,
                       const int2* tiles, int ntiles, const int* tile_off, int nsys, float eps2,
                       float G, double* partial, double* out) {
  if (ntiles <= 0) return;
  ens_energy_kernel<<<ntiles, kBlock, 0, s>>>(pos, vel, off, tiles, eps2, G, partial);
  NB_LAUNCH_CHECK();
  ens_energy_final_kernel<<<(nsys + 63) / 64, 64, 0, s>>>(partial, tile_off, nsys, out);
  NB_LAUNCH_CHECK();
}

}  // namespace nb
