// Barnes-Hut tree walks (forces_bh.cu:333-447).
//
// Targets are visited in Morton order t = 0..n-1 (pos_sorted[t]); the result
// goes to acc[perm[t]] in original order, optionally followed by the closing
// kick of the leapfrog step on vel[perm[t]].
//
// Opening test (forces_bh.cu:356-359): node accepted iff size2 < theta2 * d2,
// d2 = |com - target|^2 in the canonical FP32 order, strict, so theta = 0
// never accepts.  Children are pushed left then right (right popped first).
//
//   walk_thread_kernel   one target per lane, private stack (traverse_kernel).
//   walk_wave_kernel     group width 64 = one wave64: the whole wave shares
//                        one path; a node opens if any live lane votes open
//                        (ballot); the stack lives in two VGPRs, one entry
//                        per lane (lane-select write / readlane), so the node id is
//                        wave-uniform and node records come in through scalar
//                        loads.
//   walk_group_kernel<W> group widths 16 and 32: several groups per wave,
//                        each with its own LDS stack and a ballot restricted
//                        to its lanes; W = 32 is the reference's warp walk
//                        (traverse_warp_kernel) visit for visit.
// COUNT variants also record interactions and opening tests (per target for
// the thread walk, per group otherwise).
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

__device__ __forceinline__ void finish(int t, int n, float ax, float ay, float az, float G,
                                       const unsigned* __restrict__ perm, float4* __restrict__ acc,
                                       float4* __restrict__ vel, float half_dt) {
  float4 r = make_float4(G * ax, G * ay, G * az, 0.f);
  unsigned o = perm[t];
  if (o >= (unsigned)n) return;  // only on a corrupt permutation
  if (acc) acc[o] = r;
  if (vel) {
    float4 v = vel[o];
    v.x = fmaf(half_dt, r.x, v.x);
    v.y = fmaf(half_dt, r.y, v.y);
    v.z = fmaf(half_dt, r.z, v.z);
    vel[o] = v;
  }
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void walk_thread_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int nm1 = n - 1;
  const float4 b = pos_sorted[t];
  float ax = 0.f, ay = 0.f, az = 0.f;
  int stack[kStackDepth];
  int sp = 0, ni = 0, no = 0;
  stack[sp++] = 0;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  while (sp > 0) {
    const int id = stack[--sp];
    if ((unsigned)id >= nodes_total || ++pops > nodes_total) {
      *error = 1;
      break;
    }
    float dx, dy, dz, d2, m;
    if (id >= nm1) {
      float4 src = pos_sorted[id - nm1];
      dx = src.x - b.x; dy = src.y - b.y; dz = src.z - b.z;
      d2 = dist2(dx, dy, dz);
      m = src.w;
    } else {
      const float4 c = nodes[id].com;
      dx = c.x - b.x; dy = c.y - b.y; dz = c.z - b.z;
      d2 = dist2(dx, dy, dz);
      if (COUNT) ++no;
      if (!(nodes[id].size2 < theta2 * d2)) {
        if (sp + 2 > kStackDepth) {
          *overflow = 1;
          break;
        }
        int2 ch = *reinterpret_cast<const int2*>(&nodes[id].left);
        stack[sp++] = ch.x;
        stack[sp++] = ch.y;
        continue;
      }
      m = c.w;
    }
    accumulate(dx, dy, dz, d2, m, eps2, ax, ay, az);
    if (COUNT) ++ni;
  }
  if (COUNT) {
    counts[t] = ni;
    opens[t] = no;
  }
  finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void walk_wave_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int gbase = t - lane;
  if (gbase >= n) return;  // wave-uniform
  const bool live = t < n;
  const int nm1 = n - 1;
  const float4 b = pos_sorted[live ? t : nm1];
  float ax = 0.f, ay = 0.f, az = 0.f;
  int stk0 = 0, stk1 = 0;  // stack entry k lives in lane k&63 of stk0 (k<64) or stk1
  int sp = 1;              // entry 0 = root (0)
  int ni = 0, no = 0;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  while (sp > 0) {
    --sp;
    const int id = sp < 64 ? __builtin_amdgcn_readlane(stk0, sp)
                           : __builtin_amdgcn_readlane(stk1, sp - 64);
    if ((unsigned)id >= nodes_total || ++pops > nodes_total) {
      if (lane == 0) *error = 1;
      break;
    }
    float dx, dy, dz, d2, m;
    if (id >= nm1) {
      const float4 src = pos_sorted[id - nm1];
      dx = src.x - b.x; dy = src.y - b.y; dz = src.z - b.z;
      d2 = dist2(dx, dy, dz);
      m = src.w;
    } else {
      const float4 c = nodes[id].com;
      const float s2 = nodes[id].size2;
      dx = c.x - b.x; dy = c.y - b.y; dz = c.z - b.z;
      d2 = dist2(dx, dy, dz);
      if (COUNT) ++no;
      const bool open = live && !(s2 < theta2 * d2);
      if (__ballot(open)) {
        if (sp + 2 > kStackDepth) {
          if (lane == 0) *overflow = 1;
          break;
        }
        const int l = __builtin_amdgcn_readfirstlane(nodes[id].left);
        const int r = __builtin_amdgcn_readfirstlane(nodes[id].right);
        // lane (sp & 63) of the stack register takes the entry
        if (sp < 64) stk0 = (lane == sp) ? l : stk0;
        else stk1 = (lane == sp - 64) ? l : stk1;
        ++sp;
        if (sp < 64) stk0 = (lane == sp) ? r : stk0;
        else stk1 = (lane == sp - 64) ? r : stk1;
        ++sp;
        continue;
      }
      m = c.w;
    }
    accumulate(dx, dy, dz, d2, m, eps2, ax, ay, az);
    if (COUNT) ++ni;
  }
  if (COUNT && lane == 0) {
    counts[gbase >> 6] = ni;
    opens[gbase >> 6] = no;
  }
  if (live) finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

template <int W, bool COUNT>
__global__ __launch_bounds__(kBlock) void walk_group_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error) {
  constexpr int NG = kBlock / W;
  __shared__ int stk[NG][kStackDepth];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int g = threadIdx.x / W;
  const int glane = threadIdx.x % W;
  const int gbase = t - glane;
  if (gbase >= n) return;  // group-uniform
  const bool live = t < n;
  const int nm1 = n - 1;
  const unsigned long long gmask = ((1ull << W) - 1ull) << (lane - glane);
  const float4 b = pos_sorted[live ? t : nm1];
  float ax = 0.f, ay = 0.f, az = 0.f;
  if (glane == 0) stk[g][0] = 0;
  int sp = 1, ni = 0, no = 0;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  while (sp > 0) {
    --sp;
    const int id = stk[g][sp];
    if ((unsigned)id >= nodes_total || ++pops > nodes_total) {
      if (glane == 0) *error = 1;
      break;
    }
    float dx, dy, dz, d2, m;
    if (id >= nm1) {
      const float4 src = pos_sorted[id - nm1];
      dx = src.x - b.x; dy = src.y - b.y; dz = src.z - b.z;
      d2 = dist2(dx, dy, dz);
      m = src.w;
    } else {
      const float4 c = nodes[id].com;
      const float s2 = nodes[id].size2;
      dx = c.x - b.x; dy = c.y - b.y; dz = c.z - b.z;
      d2 = dist2(dx, dy, dz);
      if (COUNT) ++no;
      const bool open = live && !(s2 < theta2 * d2);
      if (__ballot(open) & gmask) {
        if (sp + 2 > kStackDepth) {
          if (glane == 0) *overflow = 1;
          break;
        }
        if (glane == 0) {
          int2 ch = *reinterpret_cast<const int2*>(&nodes[id].left);
          stk[g][sp] = ch.x;
          stk[g][sp + 1] = ch.y;
        }
        sp += 2;
        continue;
      }
      m = c.w;
    }
    accumulate(dx, dy, dz, d2, m, eps2, ax, ay, az);
    if (COUNT) ++ni;
  }
  if (COUNT && glane == 0) {
    counts[gbase / W] = ni;
    opens[gbase / W] = no;
  }
  if (live) finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

template <bool COUNT>
void walk_dispatch(hipStream_t s, int traverse, int width, const WalkArgs& a) {
  const int g = (a.n + kBlock - 1) / kBlock;
#define NB_WALK_ARGS                                                                          \
  a.pos_sorted, a.perm, a.n, a.nodes, a.theta2, a.eps2, a.G, a.acc, a.vel, a.half_dt, a.counts, \
      a.opens, a.overflow, a.error
  if (traverse == 0) {
    walk_thread_kernel<COUNT><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
  } else if (width == 64) {
    walk_wave_kernel<COUNT><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
  } else if (width == 32) {
    walk_group_kernel<32, COUNT><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
  } else if (width == 16) {
    walk_group_kernel<16, COUNT><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
  } else {
    throw Error(kErrInvalid, "group width must be 16, 32 or 64");
  }
#undef NB_WALK_ARGS
  NB_LAUNCH_CHECK();
}

}  // namespace

void launch_walk(hipStream_t s, int traverse, int group_width, const WalkArgs& a) {
  if (a.n <= 0) return;
  if (a.counts)
    walk_dispatch<true>(s, traverse, group_width, a);
  else
    walk_dispatch<false>(s, traverse, group_width, a);
}

}  // namespace nb
