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
// over 4096 consecutive leaves with LDS arrival counters; pass 2 (one
// workgroup) finishes the few nodes whose leaf range spans workgroups.  No
// step needs device-wide release/acquire (which would write back / invalidate
// an XCD's L2 on every climb step).  The combine is FP64 from FP32 children,
// rounded once.
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

__device__ __forceinline__ uint32_t quant21(float p, double lo, double inv) {
  double v = ((double)p - lo) * inv * 2097152.0;
  v = (v > 0.0) ? v : 0.0;
  v = (v > 2097151.0) ? 2097151.0 : v;
  return (uint32_t)v;
}

template <typename K>
__device__ __forceinline__ K make_key(float4 p, double lx, double ly, double lz, double inv) {
  uint64_t k = (spread21(quant21(p.x, lx, inv)) << 2) | (spread21(quant21(p.y, ly, inv)) << 1) |
               spread21(quant21(p.z, lz, inv));
  if (sizeof(K) == 4) return (K)(k >> 33);
  return (K)k;
}

template <typename K>
__global__ __launch_bounds__(kBlock) void morton_hist_kernel(const float4* __restrict__ pos, int n,
                                                             const QuantParams* __restrict__ qp,
                                                             K* __restrict__ keys, SortState* ss,
                                                             int passes) {
  __shared__ unsigned h[kMaxPasses][256];
  for (int p = 0; p < passes; ++p) h[p][threadIdx.x] = 0;
  __syncthreads();
  const double lx = qp->lo[0], ly = qp->lo[1], lz = qp->lo[2], inv = qp->inv;
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

template <typename K, bool FIRST, bool GATHER>
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
  const int base = (int)tile * kSortTile + w * (64 * kSortItems);

  K key[kSortItems];
  unsigned val[kSortItems], rank[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    int idx = base + j * 64 + lane;
    bool ok = idx < n;
    key[j] = ok ? kin[idx] : (K)~(K)0;
    val[j] = FIRST ? (unsigned)idx : (ok ? vin[idx] : 0u);
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
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
  for (int j = 0; j < kSortItems; ++j) {
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
                                                        int* __restrict__ parent,
                                                        unsigned* __restrict__ visit,
                                                        unsigned* __restrict__ frontier_count) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && frontier_count) *frontier_count = 0;
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
  parent[left] = i;
  parent[right] = i;
  visit[i] = 0;
  if (i == 0) parent[0] = -1;
}

// Combine the two children of internal node `node` (forces_bh.cu:293-324):
// FP64 mass-weighted centre of mass rounded once to FP32, AABB union, squared
// longest edge.  Returns false if a child id is out of range.
__device__ __forceinline__ bool combine_node(const float4* __restrict__ pos_sorted, int nm1,
                                             unsigned nodes_total, Node* nodes, float4* box_lo,
                                             float4* box_hi, int node) {
  int2 ch = *reinterpret_cast<const int2*>(&nodes[node].left);
  if ((unsigned)ch.x >= nodes_total || (unsigned)ch.y >= nodes_total) return false;
  float4 cl, ll, hl, cr, lr, hr;
  if (ch.x >= nm1) {
    cl = pos_sorted[ch.x - nm1];
    ll = hl = cl;
  } else {
    cl = nodes[ch.x].com;
    ll = box_lo[ch.x];
    hl = box_hi[ch.x];
  }
  if (ch.y >= nm1) {
    cr = pos_sorted[ch.y - nm1];
    lr = hr = cr;
  } else {
    cr = nodes[ch.y].com;
    lr = box_lo[ch.y];
    hr = box_hi[ch.y];
  }
  double ml = (double)cl.w, mr = (double)cr.w, m = ml + mr;
  float4 c;
  c.x = (float)(((double)cl.x * ml + (double)cr.x * mr) / m);
  c.y = (float)(((double)cl.y * ml + (double)cr.y * mr) / m);
  c.z = (float)(((double)cl.z * ml + (double)cr.z * mr) / m);
  c.w = (float)m;
  float4 lo = make_float4(fminf(ll.x, lr.x), fminf(ll.y, lr.y), fminf(ll.z, lr.z), 0.f);
  float4 hi = make_float4(fmaxf(hl.x, hr.x), fmaxf(hl.y, hr.y), fmaxf(hl.z, hr.z), 0.f);
  float s = fmaxf(hi.x - lo.x, fmaxf(hi.y - lo.y, hi.z - lo.z));
  nodes[node].com = c;
  nodes[node].size2 = s * s;
  box_lo[node] = lo;
  box_hi[node] = hi;
  return true;
}

// Pass 1: workgroup b owns the sorted leaves [b*L, b*L + L).  Internal node i
// covers the leaves between i and nodes[i].other, so a node whose range lies
// inside the workgroup's leaves has its id in [lo, hi] and is reached only by
// this workgroup: its arrival counter lives in LDS.  A climb that reaches a
// node spanning workgroups stops and records the finished child in
// `frontier` for pass 2.
constexpr int kComThreads = 1024;
constexpr int kComLeaves = 4096;

__global__ __launch_bounds__(kComThreads) void com_local_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int* __restrict__ parent,
    float4* box_lo, float4* box_hi, int* frontier, unsigned* frontier_count, int* error) {
  __shared__ unsigned s_visit[kComLeaves];
  for (int k = threadIdx.x; k < kComLeaves; k += blockDim.x) s_visit[k] = 0;
  __syncthreads();
  const int nm1 = n - 1;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  const int lo = blockIdx.x * kComLeaves;
  const int hi = min(n - 1, lo + kComLeaves - 1);
  for (int leaf = lo + (int)threadIdx.x; leaf <= hi; leaf += blockDim.x) {
    int child = nm1 + leaf;
    int node = parent[child];
    while (node >= 0) {
      if (node >= nm1) {
        *error = 1;
        break;
      }
      const int other = nodes[node].other;
      if (min(node, other) < lo || max(node, other) > hi) {
        frontier[atomic_add_relaxed(frontier_count, 1u)] = child;
        break;
      }
      unsigned prev = __hip_atomic_fetch_add(&s_visit[node - lo], 1u, __ATOMIC_ACQ_REL,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
      if (prev != 1u) {  // first arrival stops; a third would mean a malformed tree
        if (prev > 1u) *error = 1;
        break;
      }
      if (!combine_node(pos_sorted, nm1, nodes_total, nodes, box_lo, box_hi, node)) {
        *error = 1;
        break;
      }
      child = node;
      node = parent[node];
    }
  }
}

// Pass 2: one workgroup climbs from every frontier entry through the nodes
// that span workgroups (pass 1's results are visible across the kernel
// boundary; everything in here is one workgroup, so workgroup scope holds).
__global__ __launch_bounds__(kComThreads) void com_top_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int* __restrict__ parent,
    unsigned* visit, float4* box_lo, float4* box_hi, const int* __restrict__ frontier,
    const unsigned* frontier_count, int* error) {
  const int nm1 = n - 1;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  const unsigned count = *frontier_count;
  for (unsigned e = threadIdx.x; e < count; e += blockDim.x) {
    const int start = frontier[e];
    if ((unsigned)start >= nodes_total) {
      *error = 1;
      continue;
    }
    int node = parent[start];
    while (node >= 0) {
      if (node >= nm1) {
        *error = 1;
        break;
      }
      unsigned prev = atomic_add_wg(&visit[node], 1u);
      if (prev != 1u) {
        if (prev > 1u) *error = 1;
        break;
      }
      if (!combine_node(pos_sorted, nm1, nodes_total, nodes, box_lo, box_hi, node)) {
        *error = 1;
        break;
      }
      node = parent[node];
    }
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
  int tiles = (n + kSortTile - 1) / kSortTile;
  if (tiles > max_tiles) throw Error(kErrState, "sort look-back buffer too small");
  for (int p = 0; p < passes; ++p) {
    const K* kin = (p & 1) ? kb : ka;
    K* kout = (p & 1) ? ka : kb;
    const unsigned* vin = (p & 1) ? vb : va;
    unsigned* vout = (p & 1) ? va : vb;
    bool first = p == 0 && first_identity;
    bool gather = p == passes - 1 && gsrc != nullptr;
    if (first && gather)
      onesweep_kernel<K, true, true><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss,
                                                              lookback, max_tiles, gsrc, gdst);
    else if (first)
      onesweep_kernel<K, true, false><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss,
                                                               lookback, max_tiles, gsrc, gdst);
    else if (gather)
      onesweep_kernel<K, false, true><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss,
                                                               lookback, max_tiles, gsrc, gdst);
    else
      onesweep_kernel<K, false, false><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss,
                                                                lookback, max_tiles, gsrc, gdst);
    NB_LAUNCH_CHECK();
  }
}

}  // namespace

int sort_passes(int key_bits) { return key_bits == 30 ? 4 : 8; }

void launch_morton_hist(hipStream_t s, const float4* pos, int n, const QuantParams* qp,
                        int key_bits, void* keys, SortState* ss) {
  if (n <= 0) return;
  int g = grid_for(n, 8, 2048);
  if (key_bits == 30)
    morton_hist_kernel<uint32_t><<<g, kBlock, 0, s>>>(pos, n, qp, (uint32_t*)keys, ss, 4);
  else
    morton_hist_kernel<uint64_t><<<g, kBlock, 0, s>>>(pos, n, qp, (uint64_t*)keys, ss, 8);
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

void launch_karras(hipStream_t s, int key_bits, const void* keys, int n, Node* nodes, int* parent,
                   unsigned* visit, unsigned* frontier_count) {
  if (n < 2) return;
  int g = (n - 1 + kBlock - 1) / kBlock;
  if (key_bits == 30)
    karras_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, nodes, parent, visit,
                                                 frontier_count);
  else
    karras_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, nodes, parent, visit,
                                                 frontier_count);
  NB_LAUNCH_CHECK();
}

void launch_com(hipStream_t s, const float4* pos_sorted, int n, Node* nodes, const int* parent,
                unsigned* visit, float4* box_lo, float4* box_hi, int* frontier,
                unsigned* frontier_count, int* error) {
  if (n < 2) return;
  int g = (n + kComLeaves - 1) / kComLeaves;
  com_local_kernel<<<g, kComThreads, 0, s>>>(pos_sorted, n, nodes, parent, box_lo, box_hi,
                                             frontier, frontier_count, error);
  NB_LAUNCH_CHECK();
  if (g > 1) {
    com_top_kernel<<<1, kComThreads, 0, s>>>(pos_sorted, n, nodes, parent, visit, box_lo,
                                             box_hi, frontier, frontier_count, error);
    NB_LAUNCH_CHECK();
  }
}

}  // namespace nb
