// LBVH construction: Morton keys, onesweep radix sort, Karras radix tree,
// bottom-up centres of mass (forces_bh.cu:190-331).
//
// Morton (forces_bh.cu:192-220): the quantisation is done in FP64 from the
// FP64 origin/scale written by the bbox reduction:
//   q = clamp(trunc(((p - lo) * inv) * 2^21), 0, 2^21 - 1),
//   key63 = s(qx) << 2 | s(qy) << 1 | s(qz),   key30 = key63 >> 33.
// 30-bit keys live in 32-bit words (4 sort passes), 63-bit in 64-bit (8).
// The Morton kernel also builds the digit histograms of every pass.
//
// Sort (replaces CUB SortPairs, forces_bh.cu:490-497): LSD radix sort, 8-bit
// digits, one kernel per pass.  Each 256-thread workgroup takes a ticket-
// ordered tile of 4096 keys, ranks its keys stably in LDS (per-wave ballot
// match of equal digits), publishes per-digit tile counts and resolves its
// global offsets by decoupled look-back over earlier tiles only (earlier
// tickets are already resident, so the wait always ends), then scatters.
// Stable: ties keep input order, so equal keys end in ascending original index.
//
// Karras build (forces_bh.cu:234-277): one thread per internal node.
// COM (forces_bh.cu:281-331): leaves climb towards the root and the second
// arrival at a node combines both children.  Pass 1 does this per workgroup
// over 1024 consecutive leaves with LDS arrival counters, pass 2 per 65,536
// leaves with 2-bit LDS counters, and pass 3 (one workgroup) finishes the few
// nodes whose leaf range spans superblocks.  No
// step needs device-wide release/acquire (which would write back / invalidate
// an XCD's L2 on every climb step).  The combine is FP64 from FP32 children,
// rounded once.
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

template <typename K>
__device__ __forceinline__ K make_key(float4 p, double lx, double ly, double lz, double inv) {
  uint64_t k = (spread21(quant21(p.x, lx, inv)) << 2) | (spread21(quant21(p.y, ly, inv)) << 1) |
               spread21(quant21(p.z, lz, inv));
  if (sizeof(K) == 4) return (K)(k >> 33);
  return (K)k;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void morton_hist_kernel(const float4* __restrict__ pos, int n,
                                                             const float4* __restrict__ partial,
                                                             int nparts, QuantParams* qp,
                                                             K* __restrict__ keys, SortState* ss,
                                                             int passes) {
  __shared__ unsigned h[kMaxPasses][256];
  __shared__ float s_box[kBlock / 64][6];
  __shared__ double s_q[4];
  for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0;
  // global bounding box from the per-workgroup partials (forces_bh.cu:158-188)
  float l[3] = {3.402823466e38f, 3.402823466e38f, 3.402823466e38f};
  float u[3] = {-3.402823466e38f, -3.402823466e38f, -3.402823466e38f};
  for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
    const float4 a = partial[2 * k], b = partial[2 * k + 1];
    l[0] = fminf(l[0], a.x); l[1] = fminf(l[1], a.y); l[2] = fminf(l[2], a.z);
    u[0] = fmaxf(u[0], b.x); u[1] = fmaxf(u[1], b.y); u[2] = fmaxf(u[2], b.z);
  }
  for (int a = 0; a < 3; ++a) {
    l[a] = wave_min(l[a]);
    u[a] = wave_max(u[a]);
  }
  if ((threadIdx.x & 63) == 0)
    for (int a = 0; a < 3; ++a) {
      s_box[threadIdx.x >> 6][a] = l[a];
      s_box[threadIdx.x >> 6][3 + a] = u[a];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    // FP64 origin and cubic scale (lbvh_check.py:45-48): lo, inv = 1/max(hi-lo)
    double ext = 0.0;
    float bl[3], bh[3];
    for (int a = 0; a < 3; ++a) {
      bl[a] = s_box[0][a];
      bh[a] = s_box[0][3 + a];
      for (int w = 1; w < kBlock / 64; ++w) {
        bl[a] = fminf(bl[a], s_box[w][a]);
        bh[a] = fmaxf(bh[a], s_box[w][3 + a]);
      }
      s_q[a] = (double)bl[a];
      const double e = (double)bh[a] - (double)bl[a];
      if (e > ext) ext = e;
    }
    s_q[3] = ext > 0.0 ? 1.0 / ext : 0.0;
    if (blockIdx.x == 0) {
      for (int a = 0; a < 3; ++a) {
        qp->lo[a] = s_q[a];
        qp->box_lo[a] = bl[a];
        qp->box_hi[a] = bh[a];
      }
      qp->inv = s_q[3];
      qp->box_lo[3] = qp->box_hi[3] = 0.f;
    }
  }
  __syncthreads();
  const double lx = s_q[0], ly = s_q[1], lz = s_q[2], inv = s_q[3];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    K k = make_key<K>(pos[i], lx, ly, lz, inv);
    keys[i] = k;
    for (int p = 0; p < passes; ++p) atomicAdd(&h[p][(unsigned)(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    unsigned c = h[p][threadIdx.x];
    if (c) atomic_add_relaxed(&ss->hist[p][threadIdx.x], c);
  }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void hist_kernel(const K* __restrict__ keys, int n,
                                                      SortState* ss, int passes) {
  __shared__ unsigned h[kMaxPasses][256];
  for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    K k = keys[i];
    for (int p = 0; p < passes; ++p) atomicAdd(&h[p][(unsigned)(k >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p) {
    unsigned c = h[p][threadIdx.x];
    if (c) atomic_add_relaxed(&ss->hist[p][threadIdx.x], c);
  }
}

__global__ __launch_bounds__(kBlock) void sort_init_kernel(SortState* ss, int passes) {
  for (int p = 0; p < passes; ++p) ss->hist[p][threadIdx.x] = 0;
  if (threadIdx.x < kMaxPasses) ss->ticket[threadIdx.x] = 0;
  if (threadIdx.x == 0) ss->epoch = ss->epoch + 1;
}

constexpr unsigned long long kFlagAgg = 1ull << 32;
constexpr unsigned long long kFlagInc = 2ull << 32;

template <typename K, bool FIRST, bool GATHER, int ITEMS>
__global__ __launch_bounds__(kBlock) void onesweep_kernel(
    const K* __restrict__ kin, const unsigned* __restrict__ vin, K* __restrict__ kout,
    unsigned* __restrict__ vout, int n, int pass, SortState* ss,
    unsigned long long* lookback, int max_tiles, const float4* __restrict__ gsrc,
    float4* __restrict__ gdst) {
  constexpr int WAVES = kBlock / 64;
  __shared__ unsigned s_cnt[WAVES][256];
  __shared__ unsigned s_goff[256];
  __shared__ unsigned s_scan[WAVES];
  __shared__ unsigned s_tile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int shift = pass * 8;
  if (tid == 0) s_tile = atomic_add_relaxed(&ss->ticket[pass], 1u);
#pragma unroll
  for (int k = 0; k < WAVES; ++k) s_cnt[k][tid] = 0;
  __syncthreads();
  const unsigned tile = s_tile;
  if (tile >= gridDim.x) {  // tickets not reset: refuse rather than write out of range
    if (tid == 0) ss->error = 1;
    return;
  }
  const unsigned epoch = ss->epoch & 0x3fffffffu;
  const int base = (int)tile * (kBlock * ITEMS) + w * (64 * ITEMS);

  K key[ITEMS];
  unsigned val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    bool ok = idx < n;
    key[j] = ok ? kin[idx] : (K)~(K)0;
    val[j] = FIRST ? (unsigned)idx : (ok ? vin[idx] : 0u);
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    unsigned d = (unsigned)(key[j] >> shift) & 255u;
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      unsigned bit = (d >> b) & 1u;
      unsigned long long bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    unsigned below = (unsigned)__popcll(peers & lt);
    unsigned old = s_cnt[w][d];
    if (below == 0) s_cnt[w][d] = old + (unsigned)__popcll(peers);
    rank[j] = old + below;
  }
  __syncthreads();

  // digit `tid`: per-wave exclusive offsets and the tile total
  unsigned total = 0;
#pragma unroll
  for (int k = 0; k < WAVES; ++k) {
    unsigned c = s_cnt[k][tid];
    s_cnt[k][tid] = total;
    total += c;
  }
  // start of bucket `tid` in the output: exclusive scan of the histogram
  unsigned hv = ss->hist[pass][tid];
  unsigned inc = wave_incl_scan(hv);
  if (lane == 63) s_scan[w] = inc;
  __syncthreads();
  unsigned bucket = inc - hv;
  for (int k = 0; k < w; ++k) bucket += s_scan[k];

  // decoupled look-back over earlier tiles for digit `tid`
  unsigned long long* row = lookback + (size_t)pass * max_tiles * 256;
  const unsigned long long tag = (unsigned long long)epoch << 34;
  unsigned excl = 0;
  if (tile == 0) {
    store_agent(row + tid, tag | kFlagInc | total);
  } else {
    store_agent(row + (size_t)tile * 256 + tid, tag | kFlagAgg | total);
    int k = (int)tile - 1;
    while (true) {
      unsigned long long v = load_agent(row + (size_t)k * 256 + tid);
      unsigned long long flag = v & (3ull << 32);
      if ((unsigned)(v >> 34) != epoch || flag == 0) {
        __builtin_amdgcn_s_sleep(1);
        continue;
      }
      excl += (unsigned)v;
      if (flag == kFlagInc) break;
      --k;
    }
    store_agent(row + (size_t)tile * 256 + tid, tag | kFlagInc | (excl + total));
  }
  s_goff[tid] = bucket + excl;
  __syncthreads();

#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    if (idx < n) {
      unsigned d = (unsigned)(key[j] >> shift) & 255u;
      unsigned dest = s_goff[d] + s_cnt[w][d] + rank[j];
      if (dest >= (unsigned)n || val[j] >= (unsigned)n) {  // only on a corrupt histogram
        ss->error = 1;
        continue;
      }
      kout[dest] = key[j];
      vout[dest] = val[j];
      if (GATHER) gdst[dest] = gsrc[val[j]];
    }
  }
}

template <typename K>
__device__ __forceinline__ int clz_key(K x) {
  if (sizeof(K) == 4) return __clz((int)(uint32_t)x);
  return __clzll((long long)(uint64_t)x);
}

// delta of forces_bh.cu:234-240 on a 32- or 64-bit key word.
template <typename K>
__device__ __forceinline__ int delta(const K* __restrict__ keys, int n, int i, K ki, int j) {
  if (j < 0 || j >= n) return -1;
  K x = ki ^ keys[j];
  if (x) return clz_key<K>(x);
  return (int)(8 * sizeof(K)) + __clz(i ^ j);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void karras_kernel(const K* __restrict__ keys, int n,
                                                        Node* __restrict__ nodes,
                                                        int4* __restrict__ up,
                                                        unsigned* __restrict__ visit,
                                                        SortState* ss, int passes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (ss && blockIdx.x == 0) {
    // the sort of this step is complete: reset its state for the next one
    for (int p = 0; p < passes; ++p) ss->hist[p][threadIdx.x] = 0;
    if (threadIdx.x < kMaxPasses) ss->ticket[threadIdx.x] = 0;
    if (threadIdx.x == 0) ss->epoch = ss->epoch + 1;
  }
  if (i >= n - 1) return;
  const K ki = keys[i];
  int d = (delta(keys, n, i, ki, i + 1) - delta(keys, n, i, ki, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, ki, i - d);
  int lmax = 2;
  while (delta(keys, n, i, ki, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t > 0; t >>= 1)
    if (delta(keys, n, i, ki, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int first = min(i, j), last = max(i, j);
  const K kf = keys[first];
  int common = delta(keys, n, first, kf, last);
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    if (split + step < last && delta(keys, n, first, kf, split + step) > common) split += step;
  } while (step > 1);
  int left = (split == first) ? (n - 1 + split) : split;
  int right = (split + 1 == last) ? (n - 1 + split + 1) : split + 1;
  *reinterpret_cast<int2*>(&nodes[i].left) = make_int2(left, right);
  nodes[i].other = j;
  // each child's upward record: parent, sibling, far end of the parent's range
  up[left] = make_int4(i, right, j, 0);
  up[right] = make_int4(i, left, j, 0);
  visit[i] = 0;
  if (i == 0) up[0] = make_int4(-1, -1, j, 0);
}

// Node data a climbing thread carries: centre of mass (x, y, z, m) and AABB.
struct NodeData {
  float4 com, lo, hi;
};

__device__ __forceinline__ NodeData leaf_data(const float4* __restrict__ pos_sorted, int leaf) {
  const float4 p = pos_sorted[leaf];
  return NodeData{p, p, p};
}

__device__ __forceinline__ NodeData load_data(const float4* __restrict__ pos_sorted, int nm1,
                                              const Node* nodes, const float4* box_lo,
                                              const float4* box_hi, int id) {
  if (id >= nm1) return leaf_data(pos_sorted, id - nm1);
  return NodeData{nodes[id].com, box_lo[id], box_hi[id]};
}

// Combine two children (forces_bh.cu:313-324): FP64 mass-weighted centre of
// mass rounded once to FP32, AABB union, squared longest edge.  The formula is
// symmetric in the two children (IEEE + and fmin/fmax commute), so the caller
// may pass them in either order.
__device__ __forceinline__ NodeData combine(const NodeData& a, const NodeData& b) {
  const double ma = (double)a.com.w, mb = (double)b.com.w, m = ma + mb;
  NodeData r;
  r.com.x = (float)(((double)a.com.x * ma + (double)b.com.x * mb) / m);
  r.com.y = (float)(((double)a.com.y * ma + (double)b.com.y * mb) / m);
  r.com.z = (float)(((double)a.com.z * ma + (double)b.com.z * mb) / m);
  r.com.w = (float)m;
  r.lo = make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), 0.f);
  r.hi = make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), 0.f);
  return r;
}

__device__ __forceinline__ void store_node(Node* nodes, float4* box_lo, float4* box_hi, int id,
                                           const NodeData& d) {
  const float s = fmaxf(d.hi.x - d.lo.x, fmaxf(d.hi.y - d.lo.y, d.hi.z - d.lo.z));
  nodes[id].com = d.com;
  nodes[id].size2 = s * s;
  box_lo[id] = d.lo;
  box_hi[id] = d.hi;
}

// Frontier lists between the passes live in fixed per-workgroup segments
// (a contiguous leaf range splits into at most ~2 x depth maximal subtrees, far
// below the segment size; overflow is reported as an error), so the passes
// need no global atomics: entries are collected in LDS and one count per
// workgroup is written.
constexpr int kComThreads = 1024;
constexpr int kComLeaves = 1024;    // pass-1 leaves per workgroup (one per thread)
constexpr int kComSuper = 65536;    // pass-2 leaves per workgroup
constexpr int kSeg = 256;           // frontier entries per workgroup segment
constexpr int kTopSegs = 4096;      // superblocks the top pass can take (n <= 2^28)

__device__ __forceinline__ void seg_push(int* s_list, unsigned* s_cnt, int v, int* error) {
  const unsigned k = __hip_atomic_fetch_add(s_cnt, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
  if (k < (unsigned)kSeg) s_list[k] = v;
  else *error = 1;
}

// Pass 1: workgroup b owns the sorted leaves [b*L, b*L + L), one per thread.
// Internal node i covers the leaves between i and its `other` end, so a node
// whose range lies inside the workgroup's leaves has its id in [lo, hi] and is
// reached only by this workgroup.  The workgroup preloads the upward records
// of its leaves and of internal nodes lo..hi into LDS (coalesced), keeps the
// arrival counters and the finished nodes' COM/AABB in LDS, and so climbs
// without dependent global loads: global memory sees the leaf positions, the
// final node stores and the frontier.  A climb that reaches a node spanning
// workgroups stops and records the finished child (with its AABB, which only
// the later passes need) in the workgroup's frontier segment.
__global__ __launch_bounds__(kComThreads) void com_local_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int4* __restrict__ up,
    float4* box_lo, float4* box_hi, int* seg1, unsigned* cnt1, int* error) {
  __shared__ unsigned s_visit[kComLeaves];
  __shared__ int3 s_up_leaf[kComLeaves];
  __shared__ int3 s_up_node[kComLeaves];
  __shared__ float4 s_com[kComLeaves];
  __shared__ float3 s_lo[kComLeaves], s_hi[kComLeaves];
  __shared__ int s_front[kSeg];
  __shared__ unsigned s_nfront;
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  const int lo = blockIdx.x * kComLeaves;
  const int hi = min(n - 1, lo + kComLeaves - 1);
  const int t = threadIdx.x;
  s_visit[t] = 0;
  if (t == 0) s_nfront = 0;
  if (lo + t <= hi) {
    const int4 ul = up[nm1 + lo + t];
    s_up_leaf[t] = make_int3(ul.x, ul.y, ul.z);
    if (lo + t < nm1) {
      const int4 un = up[lo + t];
      s_up_node[t] = make_int3(un.x, un.y, un.z);
    }
  }
  __syncthreads();
  const int leaf = lo + t;
  if (leaf <= hi) {
    int child = nm1 + leaf;
    NodeData cur = leaf_data(pos_sorted, leaf);
    int3 u = s_up_leaf[t];
    while (true) {
      const int node = u.x;
      if (node < 0) break;  // reached the root
      if (node >= nm1 || (unsigned)u.y >= (unsigned)total) {
        *error = 1;
        break;
      }
      if (min(node, u.z) < lo || max(node, u.z) > hi) {
        // the finished child is a frontier entry: later passes need its AABB
        if (child < nm1) {
          box_lo[child] = cur.lo;
          box_hi[child] = cur.hi;
        }
        seg_push(s_front, &s_nfront, child, error);
        break;
      }
      const int slot = node - lo;
      const unsigned prev = __hip_atomic_fetch_add(&s_visit[slot], 1u, __ATOMIC_ACQ_REL,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
      if (prev != 1u) {  // first arrival stops; a third would mean a malformed tree
        if (prev > 1u) *error = 1;
        break;
      }
      // the sibling is a local leaf (global) or a local internal node (LDS)
      NodeData sib;
      if (u.y >= nm1) {
        sib = leaf_data(pos_sorted, u.y - nm1);
      } else {
        const int ss = u.y - lo;
        sib.com = s_com[ss];
        sib.lo = make_float4(s_lo[ss].x, s_lo[ss].y, s_lo[ss].z, 0.f);
        sib.hi = make_float4(s_hi[ss].x, s_hi[ss].y, s_hi[ss].z, 0.f);
      }
      cur = combine(cur, sib);
      s_com[slot] = cur.com;
      s_lo[slot] = make_float3(cur.lo.x, cur.lo.y, cur.lo.z);
      s_hi[slot] = make_float3(cur.hi.x, cur.hi.y, cur.hi.z);
      const float sz = fmaxf(cur.hi.x - cur.lo.x, fmaxf(cur.hi.y - cur.lo.y, cur.hi.z - cur.lo.z));
      nodes[node].com = cur.com;
      nodes[node].size2 = sz * sz;
      child = node;
      u = s_up_node[slot];
    }
  }
  __syncthreads();
  const unsigned cnt = min(s_nfront, (unsigned)kSeg);
  if (t < (int)cnt) seg1[blockIdx.x * kSeg + t] = s_front[t];
  if (t == 0) cnt1[blockIdx.x] = cnt;
}

// Climb through global node data with arrival counters in LDS, shared by
// passes 2 and 3.  `counter(node)` returns the previous arrival count.
template <typename Counter>
__device__ __forceinline__ void climb_global(const float4* __restrict__ pos_sorted, int n,
                                             Node* nodes, const int4* __restrict__ up,
                                             float4* box_lo, float4* box_hi, int child,
                                             int lo, int hi, int* s_list, unsigned* s_cnt,
                                             Counter counter, int* error) {
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  NodeData cur = load_data(pos_sorted, nm1, nodes, box_lo, box_hi, child);
  while (true) {
    const int4 u = up[child];
    const int node = u.x;
    if (node < 0) break;
    if (node >= nm1 || (unsigned)u.y >= (unsigned)total) {
      *error = 1;
      break;
    }
    if (min(node, u.z) < lo || max(node, u.z) > hi) {
      seg_push(s_list, s_cnt, child, error);
      break;
    }
    const unsigned prev = counter(node);
    if (prev != 1u) {
      if (prev > 1u) *error = 1;
      break;
    }
    cur = combine(cur, load_data(pos_sorted, nm1, nodes, box_lo, box_hi, u.y));
    store_node(nodes, box_lo, box_hi, node, cur);
    child = node;
  }
}

// Pass 2: workgroup s owns the superblock of leaves [s*S, s*S + S): it climbs
// from the frontier segments of its pass-1 workgroups with 2-bit arrival
// counters in LDS (a node arrives at most twice, so adding 1 << 2k never
// carries into the next field), indexed by node - lo.
__global__ __launch_bounds__(kComThreads) void com_mid_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int4* __restrict__ up,
    float4* box_lo, float4* box_hi, const int* __restrict__ seg1,
    const unsigned* __restrict__ cnt1, int nblocks1, int* seg2, unsigned* cnt2, int* error) {
  __shared__ unsigned s_visit[kComSuper / 16];
  __shared__ int s_front[kSeg];
  __shared__ unsigned s_nfront;
  for (int k = threadIdx.x; k < kComSuper / 16; k += blockDim.x) s_visit[k] = 0;
  if (threadIdx.x == 0) s_nfront = 0;
  __syncthreads();
  const int lo = blockIdx.x * kComSuper;
  const int hi = min(n - 1, lo + kComSuper - 1);
  constexpr int kPer = kComSuper / kComLeaves;  // pass-1 workgroups per superblock
  const int b0 = blockIdx.x * kPer;
  auto counter = [&](int node) {
    const int slot = node - lo;
    const unsigned shift = 2u * (unsigned)(slot & 15);
    const unsigned old = __hip_atomic_fetch_add(&s_visit[slot >> 4], 1u << shift,
                                                __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    return (old >> shift) & 3u;
  };
  // compact the segments' entries so that entry j goes to thread j
  __shared__ unsigned s_off[kPer + 1];
  if (threadIdx.x == 0) {
    unsigned acc = 0;
    for (int b = 0; b < kPer; ++b) {
      s_off[b] = acc;
      acc += (b0 + b < nblocks1) ? cnt1[b0 + b] : 0u;
    }
    s_off[kPer] = acc;
  }
  __syncthreads();
  for (unsigned j = threadIdx.x; j < s_off[kPer]; j += blockDim.x) {
    int b = 0;
    for (int step = kPer / 2; step > 0; step >>= 1)  // last b with s_off[b] <= j
      if (s_off[b + step] <= j) b += step;
    climb_global(pos_sorted, n, nodes, up, box_lo, box_hi, seg1[(b0 + b) * kSeg + (j - s_off[b])],
                 lo, hi, s_front, &s_nfront, counter, error);
  }
  __syncthreads();
  const unsigned cnt = min(s_nfront, (unsigned)kSeg);
  if (threadIdx.x < cnt) seg2[blockIdx.x * kSeg + threadIdx.x] = s_front[threadIdx.x];
  if (threadIdx.x == 0) cnt2[blockIdx.x] = cnt;
}

// Pass 3: one workgroup climbs from every superblock's frontier segment
// through the few nodes that span superblocks, with arrival counters in global
// memory (one workgroup, so workgroup scope holds; earlier passes are visible
// across the kernel boundaries).
__global__ __launch_bounds__(kComThreads) void com_top_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int4* __restrict__ up,
    unsigned* visit, float4* box_lo, float4* box_hi, const int* __restrict__ seg2,
    const unsigned* __restrict__ cnt2, int nblocks2, int* error) {
  __shared__ int s_dummy[1];
  __shared__ unsigned s_dummy_cnt;
  if (threadIdx.x == 0) s_dummy_cnt = 0;
  __syncthreads();
  auto counter = [&](int node) { return atomic_add_wg(&visit[node], 1u); };
  // compact the entries of all superblock segments so entry j goes to thread j
  __shared__ unsigned s_off[kTopSegs + 1];
  if (threadIdx.x == 0) {
    unsigned acc = 0;
    for (int b = 0; b < nblocks2; ++b) {
      s_off[b] = acc;
      acc += cnt2[b];
    }
    s_off[nblocks2] = acc;
  }
  __syncthreads();
  const unsigned total_entries = s_off[nblocks2];
  for (unsigned j = threadIdx.x; j < total_entries; j += blockDim.x) {
    int b = 0;
    for (int step = kTopSegs / 2; step > 0; step >>= 1)  // last b with s_off[b] <= j
      if (b + step <= nblocks2 && s_off[b + step] <= j) b += step;
    // the whole tree is in range here, so no climb stops at a frontier
    climb_global(pos_sorted, n, nodes, up, box_lo, box_hi, seg2[b * kSeg + (j - s_off[b])], 0,
                 n - 1, s_dummy, &s_dummy_cnt, counter, error);
  }
}

int grid_for(int n, int per_thread, int cap) {
  long g = ((long)n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <typename K>
void sort_typed(hipStream_t s, K* ka, K* kb, unsigned* va, unsigned* vb, int n, SortState* ss,
                unsigned long long* lookback, int max_tiles, bool first_identity,
                const float4* gsrc, float4* gdst, int passes) {
  // large inputs sort faster with 6144-key tiles (fewer tiles and look-back
  // steps; measured 4.2 vs 5.0 ms for 8 passes over 33.5M keys), smaller ones
  // with 4096-key tiles (more workgroups to fill the chip)
  const bool big = n >= (1 << 23);
  const int tile_keys = kBlock * (big ? kSortItemsBig : kSortItems);
  int tiles = (n + tile_keys - 1) / tile_keys;
  if (tiles > max_tiles) throw Error(kErrState, "sort look-back buffer too small");
  for (int p = 0; p < passes; ++p) {
    const K* kin = (p & 1) ? kb : ka;
    K* kout = (p & 1) ? ka : kb;
    const unsigned* vin = (p & 1) ? vb : va;
    unsigned* vout = (p & 1) ? va : vb;
    bool first = p == 0 && first_identity;
    bool gather = p == passes - 1 && gsrc != nullptr;
#define NB_SWEEP(F, G, I)                                                                  \
  onesweep_kernel<K, F, G, I><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss, lookback, \
                                                       max_tiles, gsrc, gdst)
    if (big) {
      if (first && gather) NB_SWEEP(true, true, kSortItemsBig);
      else if (first) NB_SWEEP(true, false, kSortItemsBig);
      else if (gather) NB_SWEEP(false, true, kSortItemsBig);
      else NB_SWEEP(false, false, kSortItemsBig);
    } else {
      if (first && gather) NB_SWEEP(true, true, kSortItems);
      else if (first) NB_SWEEP(true, false, kSortItems);
      else if (gather) NB_SWEEP(false, true, kSortItems);
      else NB_SWEEP(false, false, kSortItems);
    }
#undef NB_SWEEP
    NB_LAUNCH_CHECK();
  }
}

}  // namespace

int sort_passes(int key_bits) { return key_bits == 30 ? 4 : 8; }

void launch_morton_hist(hipStream_t s, const float4* pos, int n, const float4* partial,
                        int nparts, QuantParams* qp, int key_bits, void* keys, SortState* ss) {
  if (n <= 0) return;
  int g = grid_for(n, 8, 2048);
  if (key_bits == 30)
    morton_hist_kernel<uint32_t><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                      (uint32_t*)keys, ss, 4);
  else
    morton_hist_kernel<uint64_t><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                      (uint64_t*)keys, ss, 8);
  NB_LAUNCH_CHECK();
}

void launch_hist(hipStream_t s, const void* keys, int n, int key_bits, SortState* ss) {
  if (n <= 0) return;
  int g = grid_for(n, 8, 2048);
  if (key_bits == 30)
    hist_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, ss, 4);
  else
    hist_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, ss, 8);
  NB_LAUNCH_CHECK();
}

void launch_sort_init(hipStream_t s, SortState* ss, int passes) {
  sort_init_kernel<<<1, kBlock, 0, s>>>(ss, passes);
  NB_LAUNCH_CHECK();
}

void launch_sort(hipStream_t s, int key_bits, void* keys_a, void* keys_b, unsigned* vals_a,
                 unsigned* vals_b, int n, SortState* ss, unsigned long long* lookback,
                 int max_tiles, bool first_pass_identity, const float4* gather_src,
                 float4* gather_dst) {
  if (n <= 0) return;
  if (key_bits == 30)
    sort_typed<uint32_t>(s, (uint32_t*)keys_a, (uint32_t*)keys_b, vals_a, vals_b, n, ss, lookback,
                         max_tiles, first_pass_identity, gather_src, gather_dst, 4);
  else
    sort_typed<uint64_t>(s, (uint64_t*)keys_a, (uint64_t*)keys_b, vals_a, vals_b, n, ss, lookback,
                         max_tiles, first_pass_identity, gather_src, gather_dst, 8);
}

void launch_sort_u64(hipStream_t s, uint64_t* keys_a, uint64_t* keys_b, unsigned* vals_a,
                     unsigned* vals_b, int n, int passes, SortState* ss,
                     unsigned long long* lookback, int max_tiles, bool first_pass_identity,
                     const float4* gather_src, float4* gather_dst) {
  if (n <= 0) return;
  if (passes < 1 || passes > kMaxPasses) throw Error(kErrInvalid, "sort pass count out of range");
  sort_typed<uint64_t>(s, keys_a, keys_b, vals_a, vals_b, n, ss, lookback, max_tiles,
                       first_pass_identity, gather_src, gather_dst, passes);
}

void launch_karras(hipStream_t s, int key_bits, const void* keys, int n, Node* nodes, int4* up,
                   unsigned* visit, SortState* ss, int passes) {
  if (n < 2) return;
  int g = (n - 1 + kBlock - 1) / kBlock;
  if (key_bits == 30)
    karras_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, nodes, up, visit, ss,
                                                 passes);
  else
    karras_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, nodes, up, visit, ss,
                                                 passes);
  NB_LAUNCH_CHECK();
}

int com_scratch_ints(int n) {
  const int g1 = (n + kComLeaves - 1) / kComLeaves;
  const int g2 = (n + kComSuper - 1) / kComSuper;
  return (g1 + g2) * (kSeg + 1);
}

void launch_com(hipStream_t s, const float4* pos_sorted, int n, Node* nodes, const int4* up,
                unsigned* visit, float4* box_lo, float4* box_hi, int* scratch, int* error) {
  if (n < 2) return;
  const int g1 = (n + kComLeaves - 1) / kComLeaves;
  const int g2 = (n + kComSuper - 1) / kComSuper;
  int* seg1 = scratch;
  int* seg2 = seg1 + g1 * kSeg;
  unsigned* cnt1 = reinterpret_cast<unsigned*>(seg2 + g2 * kSeg);
  unsigned* cnt2 = cnt1 + g1;
  com_local_kernel<<<g1, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, box_lo, box_hi, seg1,
                                              cnt1, error);
  NB_LAUNCH_CHECK();
  if (g1 == 1) return;
  com_mid_kernel<<<g2, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, box_lo, box_hi, seg1, cnt1,
                                            g1, seg2, cnt2, error);
  NB_LAUNCH_CHECK();
  if (g2 == 1) return;
  if (g2 > kTopSegs) throw Error(kErrInvalid, "particle count too large for the COM top pass");
  com_top_kernel<<<1, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, visit, box_lo, box_hi, seg2,
                                           cnt2, g2, error);
  NB_LAUNCH_CHECK();
}

}  // namespace nb
