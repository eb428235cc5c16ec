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
// ordered tile of keys, ranks its keys stably in LDS (per-wave ballot match
// of equal digits), publishes per-digit tile counts and resolves its global
// offsets by decoupled look-back over earlier tiles only (earlier tickets
// are already resident, so the wait always ends).  The tile is then staged
// in LDS in sorted order and written out in that order, so consecutive
// threads write consecutive positions of each digit's run (1M 30-bit keys:
// 95.5 us for the four passes against 103.8 scattering from registers).
// Stable: ties keep input order, so equal keys end in ascending original
// index.  A 1M-key pass is one tile per CU, bound by latency: ~5 us to load
// and rank, ~9 us until inclusive prefixes have propagated through the 256
// tiles, ~2 us to write, ~3 us between launches.
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

// Digit plan of a key word: 32-bit words hold 30-bit keys, 64-bit words 63.
template <typename K>
struct Digits {
  static constexpr int kKeyBits = sizeof(K) == 4 ? 30 : 63;
  static constexpr int kBits = sort_bits(kKeyBits);
  static constexpr int kPasses = sort_passes(kKeyBits);
  static constexpr int kWords = kPasses << kBits;
  static_assert(kWords <= kHistWords && kPasses * kBits >= kKeyBits, "digit plan");
};

// LDS histogram of every pass (pass p at p << bits) for keys k.
template <typename K>
__device__ __forceinline__ void count_digits(unsigned* h, K k) {
  constexpr int B = Digits<K>::kBits;
#pragma unroll
  for (int p = 0; p < Digits<K>::kPasses; ++p)
    atomicAdd(&h[(p << B) + ((unsigned)(k >> (B * p)) & ((1u << B) - 1u))], 1u);
}

template <typename K>
__device__ __forceinline__ void flush_digits(const unsigned* h, SortState* ss) {
  for (int i = threadIdx.x; i < Digits<K>::kWords; i += kBlock) {
    const unsigned c = h[i];
    if (c) atomic_add_relaxed(&ss->hist[i], c);
  }
}

template <typename K>
__global__ __launch_bounds__(kBlock) void morton_hist_kernel(const float4* __restrict__ pos, int n,
                                                             const float4* __restrict__ partial,
                                                             int nparts, QuantParams* qp,
                                                             K* __restrict__ keys, SortState* ss) {
  __shared__ unsigned h[Digits<K>::kWords];
  __shared__ float s_box[kBlock / 64][6];
  __shared__ double s_q[4];
  for (int i = threadIdx.x; i < Digits<K>::kWords; i += kBlock) h[i] = 0;
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
  lds_barrier();  // s_q; the qp stores are for later launches
  const double lx = s_q[0], ly = s_q[1], lz = s_q[2], inv = s_q[3];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    K k = make_key<K>(pos[i], lx, ly, lz, inv);
    keys[i] = k;
    count_digits<K>(h, k);
  }
  lds_barrier();  // the histogram; the key stores are for later launches
  flush_digits<K>(h, ss);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void hist_kernel(const K* __restrict__ keys, int n,
                                                      SortState* ss) {
  __shared__ unsigned h[Digits<K>::kWords];
  for (int i = threadIdx.x; i < Digits<K>::kWords; i += kBlock) h[i] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    count_digits<K>(h, keys[i]);
  __syncthreads();
  flush_digits<K>(h, ss);
}

// Histograms and tickets to zero, a new look-back epoch (one workgroup).
__device__ __forceinline__ void reset_sort_state(SortState* ss) {
  for (int i = threadIdx.x; i < kHistWords; i += kBlock) ss->hist[i] = 0;
  if (threadIdx.x < kMaxPasses) ss->ticket[threadIdx.x] = 0;
  if (threadIdx.x == 0) ss->epoch = ss->epoch + 1;
}

__global__ __launch_bounds__(kBlock) void sort_init_kernel(SortState* ss) { reset_sort_state(ss); }

constexpr unsigned long long kFlagAgg = 1ull << 32;
constexpr unsigned long long kFlagInc = 2ull << 32;
constexpr unsigned kLookbackSpins = 1u << 20;

// D consecutive 32-bit words, moved as one LDS / global access.
template <int D>
struct alignas(4 * D) Words {
  unsigned v[D];
};

template <typename K, int BITS, bool FIRST, bool GATHER, int ITEMS>
__global__ __launch_bounds__(kBlock) void onesweep_kernel(
    const K* __restrict__ kin, const unsigned* __restrict__ vin, K* __restrict__ kout,
    unsigned* __restrict__ vout, int n, int pass, SortState* ss,
    unsigned long long* lookback, int max_tiles, const float4* __restrict__ gsrc,
    float4* __restrict__ gdst) {
  constexpr int WAVES = kBlock / 64;
  constexpr int RADIX = 1 << BITS;
  constexpr unsigned MASK = RADIX - 1;
  constexpr int TILE = kBlock * ITEMS;
  constexpr int D = RADIX / kBlock;  // consecutive digits owned by each thread
  static_assert(D * kBlock == RADIX && (D == 1 || D == 2 || D == 4), "digits per thread");
  // look-back window: earlier tiles read at once per digit (1M 30-bit keys:
  // 95.5 us for the sort reading 4, 99.0 reading 1, 97.9 reading 8)
  constexpr int LB = D == 1 ? 4 : 1;
  using W = Words<D>;
  __shared__ __align__(16) unsigned s_cnt[WAVES][RADIX];
  __shared__ __align__(16) unsigned s_off[RADIX];
  __shared__ K s_key[TILE];
  __shared__ unsigned s_val[TILE];
  __shared__ unsigned s_scan[2][WAVES];
  __shared__ unsigned s_tile;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int shift = pass * BITS;
  if (tid == 0) s_tile = atomic_add_relaxed(&ss->ticket[pass], 1u);
#pragma unroll
  for (int k = 0; k < WAVES; ++k)
#pragma unroll
    for (int e = 0; e < D; ++e) s_cnt[k][e * kBlock + tid] = 0;
  __syncthreads();
  const unsigned tile = s_tile;
  if (tile >= gridDim.x) {  // tickets not reset: refuse rather than write out of range
    if (tid == 0) ss->error = 1;
    return;
  }
  const unsigned epoch = ss->epoch & 0x3fffffffu;
  const int tile0 = (int)tile * TILE;
  const int base = tile0 + w * (64 * ITEMS);

  K key[ITEMS];
  unsigned val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    bool ok = idx < n;
    key[j] = ok ? kin[idx] : (K)~(K)0;
    val[j] = FIRST ? (unsigned)idx : (ok ? vin[idx] : 0u);
  }
  // the bucket sizes of this pass, for the output offsets below
  const W hv = reinterpret_cast<const W*>(&ss->hist[pass << BITS])[tid];
  const unsigned long long lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    unsigned d = (unsigned)(key[j] >> shift) & MASK;
    unsigned long long peers = ~0ull;
#pragma unroll
    for (int b = 0; b < BITS; ++b) {
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

  // digits tid*D .. tid*D + D-1: per-wave exclusive offsets and tile totals
  W total{};
#pragma unroll
  for (int k = 0; k < WAVES; ++k) {
    W* row = reinterpret_cast<W*>(&s_cnt[k][0]) + tid;
    W c = *row, o;
#pragma unroll
    for (int e = 0; e < D; ++e) {
      o.v[e] = total.v[e];
      total.v[e] += c.v[e];
    }
    *row = o;
  }
  // exclusive scans over the digits: where each digit starts in the tile
  // (tile totals) and in the output (the histogram)
  unsigned tsum = 0, hsum = 0;
#pragma unroll
  for (int e = 0; e < D; ++e) {
    tsum += total.v[e];
    hsum += hv.v[e];
  }
  const unsigned tinc = wave_incl_scan(tsum), hinc = wave_incl_scan(hsum);
  if (lane == 63) {
    s_scan[0][w] = tinc;
    s_scan[1][w] = hinc;
  }
  __syncthreads();
  unsigned tstart[D], bucket[D];
  tstart[0] = tinc - tsum;
  bucket[0] = hinc - hsum;
  for (int k = 0; k < w; ++k) {
    tstart[0] += s_scan[0][k];
    bucket[0] += s_scan[1][k];
  }
#pragma unroll
  for (int e = 1; e < D; ++e) {
    tstart[e] = tstart[e - 1] + total.v[e - 1];
    bucket[e] = bucket[e - 1] + hv.v[e - 1];
  }
  {
    W t;
#pragma unroll
    for (int e = 0; e < D; ++e) t.v[e] = tstart[e];
    reinterpret_cast<W*>(s_off)[tid] = t;
  }
  __syncthreads();
  // stage the tile in LDS in sorted order: digit start + wave offset + rank
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const unsigned d = (unsigned)(key[j] >> shift) & MASK;
    const unsigned at = s_off[d] + s_cnt[w][d] + rank[j];
    s_key[at] = key[j];
    s_val[at] = val[j];
  }

  // decoupled look-back over earlier tiles for this thread's digits, LB
  // tiles at a time: the window's words are loaded together, then summed
  // from the nearest tile back to the first inclusive count or to the
  // first word not yet published (read again on the next round)
  unsigned long long* row = lookback + (size_t)pass * max_tiles * RADIX + tid * D;
  const unsigned long long tag = (unsigned long long)epoch << 34;
  unsigned excl[D];
#pragma unroll
  for (int e = 0; e < D; ++e) excl[e] = 0;
  if (tile == 0) {
#pragma unroll
    for (int e = 0; e < D; ++e) store_agent(row + e, tag | kFlagInc | total.v[e]);
  } else {
    unsigned long long* mine = row + (size_t)tile * RADIX;
#pragma unroll
    for (int e = 0; e < D; ++e) store_agent(mine + e, tag | kFlagAgg | total.v[e]);
    int next[D];  // nearest earlier tile not yet summed, per digit
#pragma unroll
    for (int e = 0; e < D; ++e) next[e] = (int)tile - 1;
    unsigned pending = (1u << D) - 1u;
    // Every earlier tile publishes within microseconds of this one; a word
    // that stays stale for kLookbackSpins rounds (about a second) means the
    // sort state was not reset, so the tile reports an error and still
    // publishes an inclusive count, and no later tile waits on it forever.
    unsigned spins = 0;
    while (pending) {
      unsigned long long v[D][LB];
#pragma unroll
      for (int e = 0; e < D; ++e)
#pragma unroll
        for (int i = 0; i < LB; ++i) {
          const int t = next[e] - i;
          // words before tile 0 read as unpublished; tile 0's inclusive
          // count always ends the sum before them
          v[e][i] = ((pending >> e) & 1u) && t >= 0 ? load_agent(row + (size_t)t * RADIX + e)
                                                    : 0ull;
        }
      bool stale = false;
#pragma unroll
      for (int e = 0; e < D; ++e) {
        if (!((pending >> e) & 1u)) continue;
        bool stop = false;
        int used = 0;
#pragma unroll
        for (int i = 0; i < LB; ++i) {
          const unsigned long long x = v[e][i];
          const unsigned long long flag = x & (3ull << 32);
          const bool ready = (unsigned)(x >> 34) == epoch && flag != 0;
          if (stop) continue;
          if (!ready) {
            stale = true;
            stop = true;
            continue;
          }
          excl[e] += (unsigned)x;
          ++used;
          if (flag == kFlagInc) {
            pending &= ~(1u << e);
            stop = true;
          }
        }
        next[e] -= used;
      }
      if (stale) {
        if (++spins > kLookbackSpins) {
          ss->error = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
#pragma unroll
    for (int e = 0; e < D; ++e) store_agent(mine + e, tag | kFlagInc | (excl[e] + total.v[e]));
  }
  __syncthreads();  // every wave has staged its keys and read the tile starts
  {
    // s_off becomes output position - tile position per digit
    W g;
#pragma unroll
    for (int e = 0; e < D; ++e) g.v[e] = bucket[e] + excl[e] - tstart[e];
    reinterpret_cast<W*>(s_off)[tid] = g;
  }
  __syncthreads();

  // scatter in tile order: consecutive threads write consecutive output
  // positions within each digit's run.  The padding (all-ones keys past n)
  // ranks after every real key of the tile, so the real keys fill [0, valid).
  const int valid = min(TILE, n - tile0);
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int i = j * kBlock + tid;
    if (i < valid) {
      const K k = s_key[i];
      const unsigned x = s_val[i];
      const unsigned dest = s_off[(unsigned)(k >> shift) & MASK] + (unsigned)i;
      if (dest >= (unsigned)n || x >= (unsigned)n) {  // only on a corrupt histogram
        ss->error = 1;
        continue;
      }
      kout[dest] = k;
      vout[dest] = x;
      if (GATHER) gdst[dest] = gsrc[x];
    }
  }
}

template <typename K>
__device__ __forceinline__ int clz_key(K x) {
  if (sizeof(K) == 4) return __clz((int)(uint32_t)x);
  return __clzll((long long)(uint64_t)x);
}

// Sorted keys around a workgroup's nodes, staged in LDS: kKarrasWin keys on
// each side of the kBlock nodes.  A node's searches probe at most about twice
// its range length away, so every node whose range is shorter than the window
// runs its searches at LDS latency; the few long-range nodes near the top of
// the tree fall back to global loads for their far probes.
constexpr int kKarrasWin = 256;
constexpr int kKarrasStage = kBlock + 2 * kKarrasWin;

template <typename K>
struct KeyWindow {
  const K* __restrict__ keys;
  const K* s;
  int wlo;
  __device__ __forceinline__ K operator()(int j) const {
    const unsigned k = (unsigned)(j - wlo);
    return k < (unsigned)kKarrasStage ? s[k] : keys[j];
  }
};

// delta of forces_bh.cu:234-240 through a key window; -1 outside [0, n).
template <typename K>
__device__ __forceinline__ int delta_w(const KeyWindow<K>& key, int n, int i, K ki, int j) {
  if (j < 0 || j >= n) return -1;
  K x = ki ^ key(j);
  if (x) return clz_key<K>(x);
  return (int)(8 * sizeof(K)) + __clz(i ^ j);
}

// Karras (2012) construction, forces_bh.cu:242-279: direction from the
// neighbours' common prefixes, range end by doubling then bisection, split by
// bisection on the common-prefix length.
template <typename K>
__global__ __launch_bounds__(kBlock) void karras_kernel(const K* __restrict__ keys, int n,
                                                        Node* __restrict__ nodes,
                                                        int4* __restrict__ up,
                                                        unsigned* __restrict__ visit,
                                                        SortState* ss) {
  __shared__ K s_key[kKarrasStage];
  const int base = blockIdx.x * kBlock;
  const int wlo = base - kKarrasWin;
  for (int k = threadIdx.x; k < kKarrasStage; k += kBlock) {
    const int j = wlo + k;
    if (j >= 0 && j < n) s_key[k] = keys[j];
  }
  if (ss && blockIdx.x == 0) {
    // the sort of this step is complete: reset its state for the next one
    reset_sort_state(ss);
  }
  __syncthreads();
  const KeyWindow<K> key{keys, s_key, wlo};
  const int i = base + threadIdx.x;
  if (i >= n - 1) return;
  const K ki = key(i);
  int d = (delta_w(key, n, i, ki, i + 1) - delta_w(key, n, i, ki, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta_w(key, n, i, ki, i - d);
  int lmax = 2;
  while (delta_w(key, n, i, ki, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t > 0; t >>= 1)
    if (delta_w(key, n, i, ki, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int first = min(i, j), last = max(i, j);
  const K kf = key(first);
  int common = delta_w(key, n, first, kf, last);
  int split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    if (split + step < last && delta_w(key, n, first, kf, split + step) > common) split += step;
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

// A leaf's data is its position; an internal node's is its COM and AABB.  The
// three loads are issued unconditionally from selected addresses (a leaf reads
// its position three times), so a wave with both kinds of lanes waits for one
// round trip, not two.
__device__ __forceinline__ NodeData load_data(const float4* __restrict__ pos_sorted, int nm1,
                                              const Node* nodes, const float4* box_lo,
                                              const float4* box_hi, int id) {
  const bool leaf = id >= nm1;
  const float4* pc = leaf ? pos_sorted + (id - nm1) : &nodes[id].com;
  const float4* pl = leaf ? pc : box_lo + id;
  const float4* ph = leaf ? pc : box_hi + id;
  return NodeData{*pc, *pl, *ph};
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
// reached only by this workgroup.  The workgroup stages its leaf positions and
// the upward records of internal nodes lo..hi in LDS (coalesced) and keeps the
// arrival counters and the finished nodes' COM/AABB in LDS, so no global load
// is waited for after staging; node stores stream out unawaited.
//
// The nodes are combined in rounds, not by one climbing thread per leaf: a
// per-leaf climb keeps a wave alive while any of its lanes climbs, and on a
// two-disk tree of 2^20 leaves that left 90% of the lanes idle.  Round r holds
// the parents that became owned in round r-1 (the second arrival at a
// parent's counter owns it), packed densely into threads 0..count-1 as
// (child, sibling, parent).  A round is a chain of LDS round trips (about
// 85 ns per read, 125 ns per atomic and per barrier on MI355X), so a round
// reads an entry, then both children's data with the parent's upward record
// at once, and one barrier separates rounds (three rotating counters).  A
// thread that owns the next parent right away combines it too (kClimb more
// combines per round), which halves the rounds of a 1024-leaf block (13 ->
// 7) without bringing back the idle lanes.  The combine is symmetric in the
// two children, so the result does not depend on which child arrives second.
// A child whose parent spans workgroups is a frontier entry (with its AABB,
// which only the later passes need).
//
// Local slots: internal node lo+k is k, leaf lo+k is kComLeaves+k.  An upward
// record in LDS is (parent slot, sibling slot), with parent = kUpRoot at the
// root and kUpOut when the parent's range leaves the workgroup.
constexpr int kUpRoot = -1;
constexpr int kUpOut = -2;
static_assert(kComLeaves == 1024, "list entries pack 11-bit child slots and 10-bit parents");
// Combines a thread may chain within one round before it hands the next owned
// parent to the following round.
constexpr int kClimb = 1;

__device__ __forceinline__ int2 local_up(int4 u, int lo, int hi, int nm1, int total, int* error) {
  if (u.x < 0) return make_int2(kUpRoot, 0);
  if (u.x >= nm1 || (unsigned)u.y >= (unsigned)total) {
    *error = 1;
    return make_int2(kUpRoot, 0);
  }
  if (min(u.x, u.z) < lo || max(u.x, u.z) > hi) return make_int2(kUpOut, 0);
  // a local parent's children lie inside its range, so the sibling is local
  const int sib = u.y >= nm1 ? u.y - nm1 : u.y;
  if (sib < lo || sib > hi) {
    *error = 1;
    return make_int2(kUpRoot, 0);
  }
  return make_int2(u.x - lo, u.y >= nm1 ? kComLeaves + sib - lo : sib - lo);
}

__global__ __launch_bounds__(kComThreads) void com_local_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int4* __restrict__ up,
    float4* box_lo, float4* box_hi, int* seg1, unsigned* cnt1, int* error) {
  __shared__ float4 s_leaf[kComLeaves];
  __shared__ float4 s_com[kComLeaves];
  __shared__ float3 s_lo[kComLeaves], s_hi[kComLeaves];
  __shared__ int2 s_up[kComLeaves];              // internal slots
  __shared__ unsigned s_visit[kComLeaves / 16];  // 2-bit arrival counters
  __shared__ unsigned s_list[2][kComLeaves];     // child | sibling << 11 | parent << 22
  __shared__ unsigned s_cnt[3];
  __shared__ int s_front[kSeg];
  __shared__ unsigned s_nfront;
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  const int lo = blockIdx.x * kComLeaves;
  const int hi = min(n - 1, lo + kComLeaves - 1);
  const int t = threadIdx.x;
  if (t < kComLeaves / 16) s_visit[t] = 0;
  if (t < 3) s_cnt[t] = 0;
  if (t == 0) s_nfront = 0;
  int2 leaf_up = make_int2(kUpRoot, 0);
  if (lo + t <= hi) {
    const float4 p = pos_sorted[lo + t];
    const int4 ul = up[nm1 + lo + t];
    int4 un = make_int4(kUpRoot, 0, 0, 0);
    if (lo + t < nm1) un = up[lo + t];
    s_leaf[t] = p;
    leaf_up = local_up(ul, lo, hi, nm1, total, error);
    if (lo + t < nm1) s_up[t] = local_up(un, lo, hi, nm1, total, error);
  }
  __syncthreads();

  // A child's data is read from both layouts at once and then selected.  The
  // empty asm statements consume every value read, so the compiler can
  // neither sink the reads into branches nor wait between them.
  struct Raw {
    float4 q, m;
    float3 l, h;
  };
  auto raw = [&](int c) {
    const int k = c & (kComLeaves - 1);
    return Raw{s_leaf[k], s_com[k], s_lo[k], s_hi[k]};
  };
  auto pin = [](Raw& a) {
    asm volatile("" : "+v"(a.q.x), "+v"(a.q.y), "+v"(a.q.z), "+v"(a.q.w), "+v"(a.m.x),
                 "+v"(a.m.y), "+v"(a.m.z), "+v"(a.m.w), "+v"(a.l.x), "+v"(a.l.y), "+v"(a.l.z),
                 "+v"(a.h.x), "+v"(a.h.y), "+v"(a.h.z));
  };
  auto select = [](const Raw& a, int c) {
    const bool leaf = c >= kComLeaves;
    NodeData d;
    d.com = leaf ? a.q : a.m;
    d.lo = leaf ? a.q : make_float4(a.l.x, a.l.y, a.l.z, 0.f);
    d.hi = leaf ? a.q : make_float4(a.h.x, a.h.y, a.h.z, 0.f);
    return d;
  };
  // a finished internal node: LDS for its parent, global memory for the walk
  auto put = [&](int slot, const NodeData& d) {
    s_com[slot] = d.com;
    s_lo[slot] = make_float3(d.lo.x, d.lo.y, d.lo.z);
    s_hi[slot] = make_float3(d.hi.x, d.hi.y, d.hi.z);
    const float sz = fmaxf(d.hi.x - d.lo.x, fmaxf(d.hi.y - d.lo.y, d.hi.z - d.lo.z));
    nodes[lo + slot].com = d.com;
    nodes[lo + slot].size2 = sz * sz;
  };
  // A finished child (slot c, data d, upward record u) reports to its parent;
  // true when this thread arrived second and so owns the parent.  The arrival
  // orders the child's LDS data (release) before a sibling's read (acquire);
  // both are LDS-only, so no global store is waited for.
  auto report = [&](int c, const NodeData& d, int2 u) {
    if (u.x == kUpRoot) return false;
    if (u.x == kUpOut) {  // the parent spans workgroups: a frontier entry
      if (c < kComLeaves) {
        box_lo[lo + c] = d.lo;
        box_hi[lo + c] = d.hi;
        seg_push(s_front, &s_nfront, lo + c, error);
      } else {
        seg_push(s_front, &s_nfront, nm1 + lo + c - kComLeaves, error);
      }
      return false;
    }
    const unsigned shift = 2u * (unsigned)(u.x & 15);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    const unsigned old = __hip_atomic_fetch_add(&s_visit[u.x >> 4], 1u << shift,
                                                __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    const unsigned prev = (old >> shift) & 3u;
    if (prev > 1u) *error = 1;
    return prev == 1u;
  };
  // Climb from finished child c while this thread owns the parent, for at
  // most kClimb combines; an owned parent beyond that is appended, as
  // (child, sibling, parent), to list `list` under counter `ctr`.
  auto climb = [&](int c, NodeData d, int2 u, int ctr, int list) {
    for (int level = 0;; ++level) {
      if (!report(c, d, u)) return;
      if (level == kClimb) {
        const unsigned k = __hip_atomic_fetch_add(&s_cnt[ctr], 1u, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_WORKGROUP);
        const unsigned e = (unsigned)c | ((unsigned)u.y << 11) | ((unsigned)u.x << 22);
        if (k < (unsigned)kComLeaves) s_list[list][k] = e;
        else *error = 1;  // more owned parents than nodes: a malformed tree
        return;
      }
      Raw rq = raw(u.y);
      int2 un = s_up[u.x];
      pin(rq);
      asm volatile("" : "+v"(un.x), "+v"(un.y));
      d = combine(d, select(rq, u.y));
      c = u.x;
      u = un;
      put(c, d);
    }
  };

  // round 0: the leaves (list 1 / counter 1)
  if (lo + t <= hi) {
    const float4 q = s_leaf[t];
    climb(kComLeaves + t, NodeData{q, q, q}, leaf_up, 1, 1);
  }
  // Round r reads counter r % 3 and list r & 1, appends to counter (r+1) % 3
  // and list (r+1) & 1, and clears counter (r+2) % 3, which every thread read
  // before this round's barrier.  A subtree of L leaves is at most L - 1
  // levels high, so a valid tree ends within kComLeaves rounds; the bound only
  // keeps a malformed one finite.
  for (int r = 1;; ++r) {
    lds_barrier();
    unsigned cnt = s_cnt[r % 3];
    unsigned e = s_list[r & 1][t];  // read with the count, used only below it
    asm volatile("" : "+v"(cnt), "+v"(e));
    if (cnt == 0) break;
    if (r > kComLeaves) {
      if (t == 0) *error = 1;
      break;
    }
    if (t == 0) s_cnt[(r + 2) % 3] = 0;
    if (t < (int)cnt) {
      const int c = (int)(e & 0x7ffu), sb = (int)((e >> 11) & 0x7ffu), slot = (int)(e >> 22);
      Raw rc = raw(c), rs = raw(sb);
      int2 u = s_up[slot];
      pin(rc);
      pin(rs);
      asm volatile("" : "+v"(u.x), "+v"(u.y));
      const NodeData d = combine(select(rc, c), select(rs, sb));
      put(slot, d);
      climb(slot, d, u, (r + 1) % 3, (r + 1) & 1);
    }
  }
  const unsigned cnt = min(s_nfront, (unsigned)kSeg);
  if (t < (int)cnt) seg1[blockIdx.x * kSeg + t] = s_front[t];
  if (t == 0) cnt1[blockIdx.x] = cnt;
}

// Where a climb goes from an upward record: 0 stop at the root, 1 malformed
// record, 2 the parent leaves [lo, hi] (frontier), 3 climb to the parent.
// Every field of the record is consumed here, before any branch, so the
// record is fetched as one 16-byte load.
__device__ __forceinline__ int climb_status(int4 u, int nm1, int total, int lo, int hi) {
  const bool bad = u.x >= nm1 || (unsigned)u.y >= (unsigned)total;
  const bool out = min(u.x, u.z) < lo || max(u.x, u.z) > hi;
  return u.x < 0 ? 0 : bad ? 1 : out ? 2 : 3;
}

// Climb through global node data, shared by passes 2 and 3.  `counter(node)`
// returns the previous arrival count.  The parent's upward record was written
// by an earlier launch, so it is fetched together with the sibling's data:
// one dependent global round trip per level.
template <typename Counter>
__device__ __forceinline__ void climb_global(const float4* __restrict__ pos_sorted, int n,
                                             Node* nodes, const int4* __restrict__ up,
                                             float4* box_lo, float4* box_hi, int child,
                                             int lo, int hi, int* s_list, unsigned* s_cnt,
                                             Counter counter, int* error) {
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  NodeData cur = load_data(pos_sorted, nm1, nodes, box_lo, box_hi, child);
  int4 u = up[child];
  while (true) {
    const int st = climb_status(u, nm1, total, lo, hi);
    if (st == 0) break;
    if (st == 1) {
      *error = 1;
      break;
    }
    if (st == 2) {
      seg_push(s_list, s_cnt, child, error);
      break;
    }
    const int node = u.x;
    const unsigned prev = counter(node);
    if (prev != 1u) {
      if (prev > 1u) *error = 1;
      break;
    }
    const int4 next = up[node];
    cur = combine(cur, load_data(pos_sorted, nm1, nodes, box_lo, box_hi, u.y));
    store_node(nodes, box_lo, box_hi, node, cur);
    child = node;
    u = next;
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
  // compact the segments' entries so that entry j goes to thread j: one wave
  // loads the kPer counts together and scans them
  static_assert(kPer == 64, "the segment scan is one wave wide");
  __shared__ unsigned s_off[kPer + 1];
  if (threadIdx.x < kPer) {
    const int b = threadIdx.x;
    const unsigned c = (b0 + b < nblocks1) ? cnt1[b0 + b] : 0u;
    const unsigned incl = wave_incl_scan(c);
    s_off[b + 1] = incl;
    if (b == 0) s_off[0] = 0;
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
  // compact the entries of all superblock segments so entry j goes to thread
  // j: each thread loads kTopPer consecutive counts, then a block-wide scan
  constexpr int kTopPer = kTopSegs / kComThreads;
  constexpr int kWaves = kComThreads / 64;
  __shared__ unsigned s_off[kTopSegs + 1];
  __shared__ unsigned s_wave[kWaves];
  unsigned c[kTopPer];
  unsigned sum = 0;
  for (int k = 0; k < kTopPer; ++k) {
    const int b = threadIdx.x * kTopPer + k;
    c[k] = b < nblocks2 ? cnt2[b] : 0u;
    sum += c[k];
  }
  const unsigned incl = wave_incl_scan(sum);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 63) s_wave[wave] = incl;
  __syncthreads();
  unsigned base = incl - sum;
  for (int w = 0; w < wave; ++w) base += s_wave[w];
  for (int k = 0; k < kTopPer; ++k) {
    s_off[threadIdx.x * kTopPer + k] = base;
    base += c[k];
  }
  if (threadIdx.x == kComThreads - 1) s_off[kTopSegs] = base;
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

// Small inputs: one workgroup sorts every key in all passes.  A onesweep
// pass over a single tile is ~10 us of load, rank, publish and scatter, and
// a small system's step is eight of them plus a launch gap each; here the
// keys stay in registers and LDS between passes, a single tile's digit
// offsets are a plain scan (no look-back), and only the last pass writes to
// global memory (with the gather of pos_sorted).  Ranking is onesweep's:
// per-wave ballot match, items in the same order, so the result is the same
// stable sort.
constexpr int kSmallThreads = 1024;
template <typename K>
struct SmallSort {
  // 8 keys per thread for 64-bit keys (8192), 12 for 32-bit (12288; 16
  // spilled registers at 1024 threads).  LDS holds one copy of the keys and
  // values plus 16-bit per-wave digit counters.  Above ~8K keys one CU
  // ranking every key is no faster than the onesweep passes (12000 63-bit
  // keys: 96 us either way), and more keys per thread lengthen each wave's
  // serial ranking (4096 keys: 42 us at 8 per thread, 52 us at 12).
  static constexpr int kItems = sizeof(K) == 8 ? 8 : 12;
  static constexpr int kMax = kSmallThreads * kItems;
};

template <typename K>
__global__ __launch_bounds__(kSmallThreads) void small_sort_kernel(
    const K* __restrict__ kin, K* __restrict__ kout, unsigned* __restrict__ vout, int n,
    int passes, const float4* __restrict__ gsrc, float4* __restrict__ gdst) {
  constexpr int ITEMS = SmallSort<K>::kItems;
  constexpr int WAVES = kSmallThreads / 64;
  __shared__ K s_key[SmallSort<K>::kMax];
  __shared__ unsigned s_val[SmallSort<K>::kMax];
  __shared__ unsigned short s_cnt[WAVES][256];
  __shared__ unsigned s_goff[256];
  __shared__ unsigned s_scan[4];
  static_assert(SmallSort<K>::kMax < 65536, "16-bit counters and offsets");
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int base = w * (64 * ITEMS);
  // a wave whose slots are all padding ranks nothing: padding (all-ones
  // keys) sorts after every real key in every pass, so it only fills the
  // slots from n up and can be left out
  const bool real_wave = base < n;
  K key[ITEMS];
  unsigned val[ITEMS], rank[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; ++j) {
    const int idx = base + j * 64 + lane;
    key[j] = idx < n ? kin[idx] : (K)~(K)0;  // padding sorts last and is never written
    val[j] = (unsigned)idx;
  }
  const unsigned long long lt = (1ull << lane) - 1ull;
  for (int p = 0; p < passes; ++p) {
    const int shift = p * 8;
    for (int k = tid; k < WAVES * 256; k += kSmallThreads) (&s_cnt[0][0])[k] = 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS && real_wave; ++j) {
      const unsigned d = (unsigned)(key[j] >> shift) & 255u;
      unsigned long long peers = ~0ull;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const unsigned bit = (d >> b) & 1u;
        const unsigned long long bal = __ballot(bit);
        peers &= bit ? bal : ~bal;
      }
      const unsigned below = (unsigned)__popcll(peers & lt);
      const unsigned old = s_cnt[w][d];
      if (below == 0) s_cnt[w][d] = (unsigned short)(old + (unsigned)__popcll(peers));
      rank[j] = old + below;
    }
    __syncthreads();
    // digit `tid`: offsets of the waves within the digit, then the digits'
    // exclusive scan (threads 0..255 are four waves)
    if (tid < 256) {
      unsigned total = 0;
#pragma unroll
      for (int k = 0; k < WAVES; ++k) {
        const unsigned c = s_cnt[k][tid];
        s_cnt[k][tid] = (unsigned short)total;
        total += c;
      }
      const unsigned inc = wave_incl_scan(total);
      if (lane == 63) s_scan[w] = inc;
      s_goff[tid] = inc - total;
    }
    __syncthreads();
    if (tid < 256)
      for (int k = 0; k < w; ++k) s_goff[tid] += s_scan[k];
    __syncthreads();
    const bool last = p == passes - 1;
#pragma unroll
    for (int j = 0; j < ITEMS && real_wave; ++j) {
      const unsigned d = (unsigned)(key[j] >> shift) & 255u;
      const unsigned dest = s_goff[d] + s_cnt[w][d] + rank[j];
      if (!last) {
        s_key[dest] = key[j];
        s_val[dest] = val[j];
      } else if (val[j] < (unsigned)n && dest < (unsigned)n) {
        // padding (original index >= n, key all ones) sorts after every real
        // key in every pass, so the real keys fill [0, n)
        kout[dest] = key[j];
        vout[dest] = val[j];
        if (gdst) gdst[dest] = gsrc[val[j]];
      }
    }
    if (last) break;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS && real_wave; ++j) {
      const int idx = base + j * 64 + lane;
      // slots from n up were filled by padding, which skipped waves never
      // wrote: restore it
      key[j] = idx < n ? s_key[idx] : (K)~(K)0;
      val[j] = idx < n ? s_val[idx] : (unsigned)idx;
    }
  }
}

int grid_for(int n, int per_thread, int cap) {
  long g = ((long)n + (long)kBlock * per_thread - 1) / ((long)kBlock * per_thread);
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

template <typename K, int BITS>
void sort_typed(hipStream_t s, K* ka, K* kb, unsigned* va, unsigned* vb, int n, SortState* ss,
                unsigned long long* lookback, int max_tiles, bool first_identity,
                const float4* gsrc, float4* gdst, int passes) {
  const int items = sort_items(n);
  const int tile_keys = kBlock * items;
  int tiles = (n + tile_keys - 1) / tile_keys;
  if (tiles > max_tiles) throw Error(kErrState, "sort look-back buffer too small");
  if ((passes << BITS) > kHistWords) throw Error(kErrInvalid, "sort digit plan too large");
  for (int p = 0; p < passes; ++p) {
    const K* kin = (p & 1) ? kb : ka;
    K* kout = (p & 1) ? ka : kb;
    const unsigned* vin = (p & 1) ? vb : va;
    unsigned* vout = (p & 1) ? va : vb;
    bool first = p == 0 && first_identity;
    bool gather = p == passes - 1 && gsrc != nullptr;
#define NB_SWEEP(F, G, I)                                                                  \
  onesweep_kernel<K, BITS, F, G, I><<<tiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, p, ss, \
                                                             lookback, max_tiles, gsrc, gdst)
    if (items == kSortItemsBig) {
      if (first && gather) NB_SWEEP(true, true, kSortItemsBig);
      else if (first) NB_SWEEP(true, false, kSortItemsBig);
      else if (gather) NB_SWEEP(false, true, kSortItemsBig);
      else NB_SWEEP(false, false, kSortItemsBig);
    } else if (items == kSortItems) {
      if (first && gather) NB_SWEEP(true, true, kSortItems);
      else if (first) NB_SWEEP(true, false, kSortItems);
      else if (gather) NB_SWEEP(false, true, kSortItems);
      else NB_SWEEP(false, false, kSortItems);
    } else {
      if (first && gather) NB_SWEEP(true, true, kSortItemsSmall);
      else if (first) NB_SWEEP(true, false, kSortItemsSmall);
      else if (gather) NB_SWEEP(false, true, kSortItemsSmall);
      else NB_SWEEP(false, false, kSortItemsSmall);
    }
#undef NB_SWEEP
    NB_LAUNCH_CHECK();
  }
}

}  // namespace

void warm_lbvh() {
  hipFuncAttributes a;
  NB_HIP(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(com_local_kernel)));
}

void launch_morton_hist(hipStream_t s, const float4* pos, int n, const float4* partial,
                        int nparts, QuantParams* qp, int key_bits, void* keys, SortState* ss) {
  if (n <= 0) return;
  // Every workgroup first reduces all bbox partials and ends with its
  // histogram atomics, so at large n fewer, fuller workgroups pay less of both
  // (1M keys at 16 per thread: 512 -> 256 workgroups, 18.1 -> 16.2 us; 128
  // was slower); small systems keep 8 per thread for parallelism (at 2K-64K
  // keys 16 per thread took 10.9 us on average instead of 7.9).
  int g = grid_for(n, n >= (1 << 19) ? 16 : 8, 2048);
  if (key_bits == 30)
    morton_hist_kernel<uint32_t><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                      (uint32_t*)keys, ss);
  else
    morton_hist_kernel<uint64_t><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                      (uint64_t*)keys, ss);
  NB_LAUNCH_CHECK();
}

void launch_hist(hipStream_t s, const void* keys, int n, int key_bits, SortState* ss) {
  if (n <= 0) return;
  int g = grid_for(n, 8, 2048);
  if (key_bits == 30)
    hist_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, ss);
  else
    hist_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, ss);
  NB_LAUNCH_CHECK();
}

void launch_sort_init(hipStream_t s, SortState* ss) {
  sort_init_kernel<<<1, kBlock, 0, s>>>(ss);
  NB_LAUNCH_CHECK();
}

void launch_sort(hipStream_t s, int key_bits, void* keys_a, void* keys_b, unsigned* vals_a,
                 unsigned* vals_b, int n, SortState* ss, unsigned long long* lookback,
                 int max_tiles, bool first_pass_identity, const float4* gather_src,
                 float4* gather_dst) {
  if (n <= 0) return;
  const void* in = sort_input(key_bits, keys_a, keys_b);
  // small inputs (identity values): one workgroup sorts on 8-bit digits in
  // all passes (4 for 30-bit keys, 8 for 63-bit), result in the a buffers
  if (first_pass_identity && key_bits == 30 && n <= SmallSort<uint32_t>::kMax) {
    small_sort_kernel<uint32_t><<<1, kSmallThreads, 0, s>>>(
        (const uint32_t*)in, (uint32_t*)keys_a, vals_a, n, 4, gather_src, gather_dst);
    NB_LAUNCH_CHECK();
    return;
  }
  if (first_pass_identity && key_bits != 30 && n <= SmallSort<uint64_t>::kMax) {
    small_sort_kernel<uint64_t><<<1, kSmallThreads, 0, s>>>(
        (const uint64_t*)in, (uint64_t*)keys_a, vals_a, n, 8, gather_src, gather_dst);
    NB_LAUNCH_CHECK();
    return;
  }
  // an odd pass count starts in the b buffers and ends in a
  const bool odd = sort_passes(key_bits) & 1;
  if (odd && !first_pass_identity)
    throw Error(kErrInvalid, "an odd-pass sort takes identity values");
  if (key_bits == 30) {
    uint32_t* a = (uint32_t*)keys_a;
    uint32_t* b = (uint32_t*)keys_b;
    sort_typed<uint32_t, sort_bits(30)>(s, odd ? b : a, odd ? a : b, odd ? vals_b : vals_a,
                                        odd ? vals_a : vals_b, n, ss, lookback, max_tiles,
                                        first_pass_identity, gather_src, gather_dst,
                                        sort_passes(30));
  } else {
    uint64_t* a = (uint64_t*)keys_a;
    uint64_t* b = (uint64_t*)keys_b;
    sort_typed<uint64_t, sort_bits(63)>(s, odd ? b : a, odd ? a : b, odd ? vals_b : vals_a,
                                        odd ? vals_a : vals_b, n, ss, lookback, max_tiles,
                                        first_pass_identity, gather_src, gather_dst,
                                        sort_passes(63));
  }
}

void launch_sort_u64(hipStream_t s, uint64_t* keys_a, uint64_t* keys_b, unsigned* vals_a,
                     unsigned* vals_b, int n, int passes, SortState* ss,
                     unsigned long long* lookback, int max_tiles, bool first_pass_identity,
                     const float4* gather_src, float4* gather_dst) {
  if (n <= 0) return;
  if (passes < 1 || passes > kMaxPasses) throw Error(kErrInvalid, "sort pass count out of range");
  sort_typed<uint64_t, kSortBits>(s, keys_a, keys_b, vals_a, vals_b, n, ss, lookback, max_tiles,
                                  first_pass_identity, gather_src, gather_dst, passes);
}

void launch_karras(hipStream_t s, int key_bits, const void* keys, int n, Node* nodes, int4* up,
                   unsigned* visit, SortState* ss) {
  if (n < 2) return;
  int g = (n - 1 + kBlock - 1) / kBlock;
  if (key_bits == 30)
    karras_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, nodes, up, visit, ss);
  else
    karras_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, nodes, up, visit, ss);
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
