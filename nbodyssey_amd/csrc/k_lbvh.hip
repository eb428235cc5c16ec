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
// of equal digits), publishes its per-digit tile counts at once and resolves
// its global offsets from earlier tiles only (earlier tickets are already
// resident, so the wait always ends): up to 1024 tiles by a two-level sum
// (the earlier tiles of its group of ~sqrt(tiles), then the earlier groups'
// totals), beyond that by decoupled look-back.  The tile is then staged in
// LDS in sorted order and written out in that order, so consecutive threads
// write consecutive positions of each digit's run.  Stable: ties keep input
// order, so equal keys end in ascending original index.  A 1M-key pass is one
// tile per CU, bound by latency: ~3 us until the last workgroup starts, ~5 us
// to load and rank, ~1.4 us to stage, ~4 us until the prefix is summed, ~2 us
// to write, ~3 us between launches (1M 30-bit keys: 76 us for the four passes;
// 95.5 us with the look-back chain, which took ~8.7 us per pass).
//
// Karras build (forces_bh.cu:234-277): one thread per internal node.
// COM (forces_bh.cu:281-331): leaves climb towards the root and the second
// arrival at a node combines both children.  Pass 1 does this per workgroup
// over 1024 consecutive leaves in LDS, pass 2 per 32,768 leaves and pass 3
// (one workgroup) over the few nodes whose leaf range spans superblocks.  The
// Karras kernel marks each internal node with the pass that owns it; passes
// 2 and 3 stage their nodes in LDS with links to their parents, let every
// frontier entry of the previous pass arrive at once, and then climb from the
// nodes that became complete with four lanes per climb (lane c computes
// component c) in a few waves.  Their critical path is the height of the
// pass's part of the tree (about 24 and 28 levels at 1M two-disk leaves) at
// ~0.35-0.5 us per level: the latency of one wave's step (one LDS read round
// trip, the FP64 combine, the slot write and the arrival count).  No step
// needs device-wide release/acquire (which would write back / invalidate an
// XCD's L2 on every climb step), and no climb step waits for a global store.
// The combine is FP64 from FP32 children, rounded once.
#include "nb_device.hpp"
#include "nb_internal.hpp"

namespace nb {
namespace {

template <typename K>
__device__ __forceinline__ K make_key(float4 p, double lx, double ly, double lz, double inv) {
  const uint32_t qx = quant21(p.x, lx, inv), qy = quant21(p.y, ly, inv), qz = quant21(p.z, lz, inv);
  if (sizeof(K) == 4)  // the 63-bit key's top 30 bits, interleaved in 32-bit words
    return (K)((spread10(qx >> 11) << 2) | (spread10(qy >> 11) << 1) | spread10(qz >> 11));
  return (K)((spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz));
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

// PER keys per thread and pass of the grid (i, i + stride, ...): their
// positions are loaded before the box reduction, which they do not need, and
// all together, so the wave waits for one round trip instead of PER.
template <typename K, int PER>
__global__ __launch_bounds__(kBlock) void morton_hist_kernel(const float4* __restrict__ pos, int n,
                                                             const float4* __restrict__ partial,
                                                             int nparts, QuantParams* qp,
                                                             K* __restrict__ keys, SortState* ss) {
  __shared__ unsigned h[Digits<K>::kWords];
  __shared__ float s_box[kBlock / 64][6];
  __shared__ double s_q[4];
  const int stride = gridDim.x * blockDim.x;
  int i0 = blockIdx.x * blockDim.x + threadIdx.x;
  float4 p[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) p[j] = pos[min(i0 + j * stride, n - 1)];
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
  for (;;) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = i0 + j * stride;
      if (i < n) {
        const K k = make_key<K>(p[j], lx, ly, lz, inv);
        keys[i] = k;
        count_digits<K>(h, k);
      }
    }
    i0 += PER * stride;
    if (i0 >= n) break;  // grids cover n in one pass up to 2048 x 256 x PER keys
#pragma unroll
    for (int j = 0; j < PER; ++j) p[j] = pos[min(i0 + j * stride, n - 1)];
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

// Sum of the `cnt` (< 64) published aggregates at p, p + RADIX, ... (D
// consecutive digits each), polling those not yet published.  A round issues
// its loads 16 at a time before it reads any of them.  A word that stays
// stale for kLookbackSpins rounds (about a second) means the sort state was
// not reset: the tile reports an error and goes on, so no tile waits forever.
template <int D, int RADIX>
__device__ __forceinline__ void sum_published(unsigned long long* p, int cnt, unsigned (&sum)[D],
                                              unsigned epoch, SortState* ss) {
  unsigned long long pending[D];
#pragma unroll
  for (int e = 0; e < D; ++e) pending[e] = cnt >= 64 ? ~0ull : (1ull << cnt) - 1ull;
  for (unsigned spins = 0;; ++spins) {
    for (int j0 = 0; j0 < cnt; j0 += 16) {
      unsigned long long v[16][D];
#pragma unroll
      for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int e = 0; e < D; ++e)
          v[i][e] = (j0 + i < cnt && ((pending[e] >> (j0 + i)) & 1ull))
                        ? load_agent(p + (size_t)(j0 + i) * RADIX + e)
                        : 0ull;
#pragma unroll
      for (int i = 0; i < 16; ++i)
#pragma unroll
        for (int e = 0; e < D; ++e) {
          const unsigned long long x = v[i][e];
          if ((unsigned)(x >> 34) == epoch && (x & (3ull << 32)) != 0) {
            sum[e] += (unsigned)x;
            pending[e] &= ~(1ull << (j0 + i));
          }
        }
    }
    unsigned long long any = 0;
#pragma unroll
    for (int e = 0; e < D; ++e) any |= pending[e];
    if (!any) break;
    if (spins >= kLookbackSpins) {
      ss->error = 1;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// Tiles up to which the onesweep pass takes the two-level prefix below
// instead of the decoupled look-back chain (1024 tiles: 4M keys).
constexpr int kGroupPrefixTiles = 1024;

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
  // look-back window of the chain: earlier tiles read at once per digit
  // (1M 30-bit keys: 95.5 us for the sort reading 4, 99.0 reading 1, 97.9
  // reading 8)
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
  // Publish this tile's digit totals at once, so that the scans and the
  // staging below overlap the other tiles' wait for them.  Per pass the
  // look-back area holds max_tiles rows of tile totals and max_tiles rows of
  // group totals, all words epoch-tagged (never cleared).
  unsigned long long* lbrow = lookback + (size_t)pass * 2 * max_tiles * RADIX + tid * D;
  const unsigned long long tag = (unsigned long long)epoch << 34;
  const int ntiles = (int)gridDim.x;
  const bool grouped = ntiles <= kGroupPrefixTiles;
#pragma unroll
  for (int e = 0; e < D; ++e)
    store_agent(lbrow + (size_t)tile * RADIX + e,
                tag | (grouped || tile != 0 ? kFlagAgg : kFlagInc) | total.v[e]);
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

  // Exclusive prefix of this tile's digit counts over the earlier tiles.
  unsigned long long* row = lbrow;
  unsigned excl[D];
#pragma unroll
  for (int e = 0; e < D; ++e) excl[e] = 0;
  if (grouped) {
    // Two levels: tiles in groups of G = ceil(sqrt(tiles)).  A tile sums
    // the aggregates of its group's earlier tiles; the group's last tile
    // publishes the group total; every tile sums the earlier groups'
    // totals.  Each level's words are loaded together, so the prefix takes a
    // few memory round trips where the chain below waits on every earlier
    // tile in turn (1M keys, 256 tiles: ~8.7 us per pass).  A tile waits only
    // on tiles with earlier tickets.
    int G = 1;
    while (G * G < ntiles) ++G;
    const int g = (int)tile / G, first = g * G;
    unsigned local[D];
#pragma unroll
    for (int e = 0; e < D; ++e) local[e] = 0;
    sum_published<D, RADIX>(row + (size_t)first * RADIX, (int)tile - first, local, epoch, ss);
    unsigned long long* groups = row + (size_t)max_tiles * RADIX;
    if ((int)tile == first + G - 1 && (int)tile < ntiles - 1)
#pragma unroll
      for (int e = 0; e < D; ++e)
        store_agent(groups + (size_t)g * RADIX + e, tag | kFlagAgg | (local[e] + total.v[e]));
    sum_published<D, RADIX>(groups, g, excl, epoch, ss);
#pragma unroll
    for (int e = 0; e < D; ++e) excl[e] += local[e];
    // More tiles: decoupled look-back over earlier tiles for this thread's
    // digits, LB tiles at a time: the window's words are loaded together,
    // then summed from the nearest tile back to the first inclusive count or
    // to the first word not yet published (read again on the next round).
  } else if (tile != 0) {  // tile 0 published its inclusive counts above
    unsigned long long* mine = row + (size_t)tile * RADIX;
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
  const lds<K>* s;
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
                                                        int2* __restrict__ up,
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
  const KeyWindow<K> key{keys, (const lds<K>*)s_key, wlo};
  const int lane = threadIdx.x & 63;
  const int i = base + threadIdx.x;
  const bool active = i < n - 1;
  // node i's record, children's upward records and pass word
  auto emit = [&](int node, int j, int first, int last, int split) {
    const int left = (split == first) ? (n - 1 + split) : split;
    const int right = (split + 1 == last) ? (n - 1 + split + 1) : split + 1;
    *reinterpret_cast<int2*>(&nodes[node].left) = make_int2(left, right);
    nodes[node].other = j;
    // each child's upward record: parent with the parent's pass word, sibling
    const unsigned vw = com_visit_word(first, last);
    up[left] = make_up(node, vw, right);
    up[right] = make_up(node, vw, left);
    visit[node] = vw;
    if (node == 0) up[0] = up_none();
  };
  // Each lane searches its node's range while it stays within kKarrasWin of
  // i (probes in the LDS window).  A longer range would make the whole wave
  // wait for that lane's chain of global probes (three searches of log2 of
  // the range each), so such nodes are marked and resolved afterwards by the
  // whole wave, 64 probes per round trip (below).
  int d = 1, dmin = 0;
  K ki = 0;
  bool is_long = false;
  if (active) {
    ki = key(i);
    d = (delta_w(key, n, i, ki, i + 1) - delta_w(key, n, i, ki, i - 1)) >= 0 ? 1 : -1;
    dmin = delta_w(key, n, i, ki, i - d);
    int lmax = 2;
    while (lmax <= kKarrasWin && delta_w(key, n, i, ki, i + lmax * d) > dmin) lmax <<= 1;
    if (lmax > kKarrasWin) {
      is_long = true;
    } else {
      int l = 0;
      for (int t = lmax >> 1; t > 0; t >>= 1)
        if (delta_w(key, n, i, ki, i + (l + t) * d) > dmin) l += t;
      const int j = i + l * d;
      const int first = min(i, j), last = max(i, j);
      const K kf = key(first);
      const int common = delta_w(key, n, first, kf, last);
      int split = first, step = last - first;
      do {
        step = (step + 1) >> 1;
        if (split + step < last && delta_w(key, n, first, kf, split + step) > common)
          split += step;
      } while (step > 1);
      emit(i, j, first, last, split);
    }
  }
  // Long ranges, one node at a time for the wave.  Both searches look for
  // the last position of a monotone predicate between a known true lo and a
  // known false hi: lane k probes lo + (k + 1) * step, and the count of true
  // probes narrows [lo, hi) 64-fold per round trip.  The results are those
  // of the doubling and bisection above (same predicate, same bounds).
  auto last_true = [&](int lo, int hi, int from, K kfrom, int dir, int thresh) {
    while (hi - lo > 1) {
      const int step = (hi - lo + 63) / 64;
      const int off = lo + (lane + 1) * step;
      const bool tv = off < hi && delta_w(key, n, from, kfrom, from + off * dir) > thresh;
      lo += __popcll(__builtin_amdgcn_ballot_w64(tv)) * step;
      hi = min(hi, lo + step);
    }
    return lo;
  };
  for (unsigned long long lm = __builtin_amdgcn_ballot_w64(is_long); lm; lm &= lm - 1) {
    const int src = __builtin_ctzll(lm);
    const int li = __builtin_amdgcn_readlane(i, src);
    const int ld = __builtin_amdgcn_readlane(d, src);
    const int lmin = __builtin_amdgcn_readlane(dmin, src);
    K lk;
    if (sizeof(K) == 4) {
      lk = (K)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)ki, src);
    } else {
      const uint64_t k64 = (uint64_t)ki;
      lk = (K)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(k64 >> 32), src)
                << 32) |
               (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)k64, src));
    }
    // range end: delta(li + l d) > lmin holds at l = kKarrasWin; bracket it
    // between kKarrasWin << (m - 1) and kKarrasWin << m (lane k probes
    // kKarrasWin << k), then narrow
    const long long far = (long long)kKarrasWin << lane;
    const bool bt = lane < 24 && far <= (long long)n &&
                    delta_w(key, n, li, lk, li + (int)far * ld) > lmin;
    const int m = __popcll(__builtin_amdgcn_ballot_w64(bt));
    const int l = last_true(kKarrasWin << (m - 1), (int)min((long long)n, (long long)kKarrasWin << m),
                            li, lk, ld, lmin);
    const int j = li + l * ld;
    const int first = min(li, j), last = max(li, j);
    const K kf = key(first);
    const int common = delta_w(key, n, first, kf, last);
    const int split = first + last_true(0, last - first, first, kf, 1, common);
    if (lane == src) emit(li, j, first, last, split);
  }
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

// n / m for the three COM components, each the IEEE FP64 quotient: one
// reciprocal of m, refined by two Newton steps, is shared by the three, and
// each quotient is q0 = n r corrected once by the residual, q = q0 + (n - m q0) r
// (Markstein's method; the same steps as the compiler's FP64 division when its
// exponent scaling does nothing, which holds for all finite operands formed
// from FP32 values: |n| < 2^257, 2^-149 <= |m| < 2^129, so every quotient and
// residual stays far inside the normal FP64 range).  A zero, infinite or NaN
// operand takes the plain division.
__device__ __forceinline__ bool finite64(double x) {
  return __builtin_fabs(x) <= 1.7976931348623157e308;
}
__device__ __forceinline__ bool nr_div_ok(double n, double m) {
  return m != 0.0 && finite64(m) && finite64(n);
}
__device__ __forceinline__ double nr_recip(double m) {
  double r = __builtin_amdgcn_rcp(m);
  r = __builtin_fma(r, __builtin_fma(-m, r, 1.0), r);
  return __builtin_fma(r, __builtin_fma(-m, r, 1.0), r);
}
__device__ __forceinline__ double nr_quot(double n, double m, double r) {
  const double q0 = n * r;  // a zero n: q0 is the correctly signed zero
  return n == 0.0 ? q0 : __builtin_fma(__builtin_fma(-m, q0, n), r, q0);
}
__device__ __forceinline__ double3 div3(double nx, double ny, double nz, double m) {
  if (!(nr_div_ok(nx, m) && finite64(ny) && finite64(nz)))
    return make_double3(nx / m, ny / m, nz / m);
  const double r = nr_recip(m);
  return make_double3(nr_quot(nx, m, r), nr_quot(ny, m, r), nr_quot(nz, m, r));
}

// Combine two children (forces_bh.cu:313-324): FP64 mass-weighted centre of
// mass rounded once to FP32, AABB union, squared longest edge.  The formula is
// symmetric in the two children (IEEE + and fmin/fmax commute), so the caller
// may pass them in either order.
__device__ __forceinline__ NodeData combine(const NodeData& a, const NodeData& b) {
  const double ma = (double)a.com.w, mb = (double)b.com.w, m = ma + mb;
  const double3 q = div3((double)a.com.x * ma + (double)b.com.x * mb,
                         (double)a.com.y * ma + (double)b.com.y * mb,
                         (double)a.com.z * ma + (double)b.com.z * mb, m);
  NodeData r;
  r.com = make_float4((float)q.x, (float)q.y, (float)q.z, (float)m);
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

// Hand-offs between the passes live in fixed per-workgroup segments of kSeg
// slots in global memory, each with one count per workgroup, so the passes
// need no global atomics and the next pass reads a segment in one round trip.
// A contiguous leaf range meets at most ~2 x depth nodes that leave it, far
// below kSeg; overflow is reported as an error.  Two kinds:
//  * FrontEntry: a finished node or leaf whose parent belongs to a later pass,
//    with its COM, AABB, link and id, so the next pass needs no further load;
//  * node records (pass 1): the internal nodes of the block's id range that a
//    later pass owns, as (id | pass << 30, link), so the next pass stages its
//    nodes without scanning the pass words.
// A link is parent id | side << 30 (side 0: the smaller id of the two
// children; a leaf's id is n - 1 + its index), or -1 at the root.
constexpr int kComThreads = 1024;  // one thread per pass-1 leaf
constexpr int kTopSegs = 8192;      // superblocks the top pass can take (n <= 2^28)
constexpr int kSideShift = 30;
constexpr int kIdMask = (1 << kSideShift) - 1;
struct FrontEntry {
  float4 com;
  float4 lo;  // w: the link, as int bits
  float4 hi;  // w: the node's id, as int bits
};
__device__ __forceinline__ int make_link(int parent, int id, int sibling) {
  return parent < 0 ? -1 : parent | ((id < sibling ? 0 : 1) << kSideShift);
}

__device__ __forceinline__ void push_entry(FrontEntry* seg, lds<unsigned>* s_cnt,
                                           const NodeData& d, int link, int id, int* error) {
  const unsigned k = __hip_atomic_fetch_add(s_cnt, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
  if (k < (unsigned)kSeg)
    seg[k] = FrontEntry{d.com, make_float4(d.lo.x, d.lo.y, d.lo.z, __int_as_float(link)),
                        make_float4(d.hi.x, d.hi.y, d.hi.z, __int_as_float(id))};
  else
    *error = 1;
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
// A child whose parent spans workgroups is a frontier entry; the internal
// nodes of [lo, hi] that a later pass owns are listed as node records.
//
// Local slots: internal node lo+k is k, leaf lo+k is kComLeaves+k.  An upward
// record in LDS is (parent slot, sibling slot), with parent = kUpRoot at the
// root, and (kUpOut, link) when the parent's range leaves the workgroup.
constexpr int kUpRoot = -1;
constexpr int kUpOut = -2;
static_assert(kComLeaves == 1024, "list entries pack 11-bit child slots and 10-bit parents");
// Combines a thread may chain within one round before it hands the next owned
// parent to the following round.
constexpr int kClimb = 1;

__device__ __forceinline__ int2 local_up(int2 u, int id, int lo, int hi, int nm1, int total,
                                         int* error) {
  const int parent = up_parent(u);
  if (parent < 0) return make_int2(kUpRoot, 0);
  if (parent >= nm1 || (unsigned)u.y >= (unsigned)total) {
    *error = 1;
    return make_int2(kUpRoot, 0);
  }
  // pass 0: the parent's leaf range lies in one pass-1 block, this child's
  if (up_pass(u) != 0) return make_int2(kUpOut, make_link(parent, id, u.y));
  // a local parent's children lie inside its range, so both are local
  const int sib = u.y >= nm1 ? u.y - nm1 : u.y;
  if (parent < lo || parent > hi || sib < lo || sib > hi) {
    *error = 1;
    return make_int2(kUpRoot, 0);
  }
  return make_int2(parent - lo, u.y >= nm1 ? kComLeaves + sib - lo : sib - lo);
}

__global__ __launch_bounds__(kComThreads) void com_local_kernel(
    float4* __restrict__ pos_sorted, int n, Node* nodes, const int2* __restrict__ up,
    const unsigned* __restrict__ visit, float4* box_lo, float4* box_hi, FrontEntry* seg1,
    unsigned* cnt1, int2* rec1, unsigned* rcnt1, const unsigned* __restrict__ perm,
    const float4* __restrict__ gsrc, unsigned* top_count, int* error) {
  __shared__ float4 s_leaf[kComLeaves];
  __shared__ float4 s_com[kComLeaves];
  __shared__ float3 s_lo[kComLeaves], s_hi[kComLeaves];
  __shared__ int2 s_up[kComLeaves];              // internal slots
  __shared__ unsigned s_visit[kComLeaves / 16];  // 2-bit arrival counters
  __shared__ unsigned s_list[2][kComLeaves];     // child | sibling << 11 | parent << 22
  __shared__ unsigned s_cnt[3];
  __shared__ unsigned s_nfront, s_nrec;
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  const int lo = blockIdx.x * kComLeaves;
  const int hi = min(n - 1, lo + kComLeaves - 1);
  const int t = threadIdx.x;
  FrontEntry* front = seg1 + (size_t)blockIdx.x * kSeg;
  if (t < kComLeaves / 16) s_visit[t] = 0;
  if (t < 3) s_cnt[t] = 0;
  // pass 2 of this build appends to the pass-3 list; the last build's pass 3
  // has read its count
  if (blockIdx.x == 0 && t == 0) *top_count = 0;
  if (t == 0) {
    s_nfront = 0;
    s_nrec = 0;
  }
  int2 leaf_up = make_int2(kUpRoot, 0);
  bool has_rec = false;  // this thread's internal node belongs to a later pass
  int2 rec = make_int2(0, 0);
  if (lo + t <= hi) {
    // with gsrc this pass also gathers the sorted positions (pos_sorted is
    // written here, read by the later passes and the walk)
    const float4 p = gsrc ? gsrc[perm[lo + t]] : pos_sorted[lo + t];
    if (gsrc) pos_sorted[lo + t] = p;
    const int2 ul = up[nm1 + lo + t];
    int2 un = up_none();
    unsigned pass = 0;
    if (lo + t < nm1) {
      un = up[lo + t];
      pass = visit[lo + t] >> 30;
    }
    s_leaf[t] = p;
    leaf_up = local_up(ul, nm1 + lo + t, lo, hi, nm1, total, error);
    if (lo + t < nm1) s_up[t] = local_up(un, lo + t, lo, hi, nm1, total, error);
    if (pass == 1 || pass == 2) {  // 3: a node that gets no COM (ensemble forests)
      has_rec = true;
      rec = make_int2((int)((unsigned)(lo + t) | (pass << 30)),
                      make_link(up_parent(un), lo + t, un.y));
    }
  }
  __syncthreads();
  if (has_rec) {
    const unsigned k = __hip_atomic_fetch_add(&s_nrec, 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k < (unsigned)kSeg) rec1[(size_t)blockIdx.x * kSeg + k] = rec;
    else *error = 1;
  }

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
      const int id = c < kComLeaves ? lo + c : nm1 + lo + c - kComLeaves;
      if (c < kComLeaves) {  // for the passes' fallback climb, which reads global memory
        box_lo[id] = d.lo;
        box_hi[id] = d.hi;
      }
      push_entry(front, (lds<unsigned>*)&s_nfront, d, u.y, id, error);
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
  if (t == 0) {  // the last round's pushes came before its barrier
    cnt1[blockIdx.x] = min(s_nfront, (unsigned)kSeg);
    rcnt1[blockIdx.x] = min(s_nrec, (unsigned)kSeg);
  }
}

// Where a climb goes from an upward record: 0 stop at the root, 1 malformed
// record, 2 the parent's pass is above max_pass (its leaf range leaves this
// pass's superblock: a frontier entry), 3 climb to the parent.  Both words of
// the record are consumed here, before any branch, so it is fetched as one
// 8-byte load.
__device__ __forceinline__ int climb_status(int2 u, int nm1, int total, unsigned max_pass) {
  const bool bad = up_parent(u) >= nm1 || (unsigned)u.y >= (unsigned)total;
  const bool out = up_pass(u) > max_pass;
  return up_root(u) ? 0 : bad ? 1 : out ? 2 : 3;
}

// Climb shared by passes 2 and 3.  The two children of a node arrive through
// its arrival word: each arriving child first puts its data (COM, AABB) in an
// LDS slot of the workgroup and then ORs the slot's tag into the word; the
// child that finds the word empty stops, and the one that finds a tag reads
// its sibling from that slot, combines, stores the node for the walk (not
// waited for) and climbs on.  A child's data thus reaches its sibling through
// LDS: no level waits for a global store to complete or for a global load of
// data written in this launch.  When the slot table is full, the child's data
// is published in global memory instead (its stores completed before the
// arrival, tag kTagGlobal), and the sibling reads it from there.
//
// `arrive(node, tag)` ORs tag into the node's word and returns the previous
// tag (0: none).  Upward records were written by an earlier launch.
struct ComSlots {
  lds<float4>* com;
  lds<float3>* lo;
  lds<float3>* hi;
  lds<unsigned>* used;  // slots handed out
  unsigned cap;
};
constexpr unsigned kTagGlobal = 0xffffu;
// A valid tree is at most 2^30 leaves deep along index bits plus the 63 key
// bits: a climb longer than this is a malformed tree.
constexpr int kMaxClimb = 128;

template <typename Arrive>
__device__ __forceinline__ void climb_exchange(const float4* __restrict__ pos_sorted, int n,
                                               Node* nodes, const int2* __restrict__ up,
                                               float4* box_lo, float4* box_hi, int child,
                                               unsigned max_pass, FrontEntry* seg,
                                               lds<unsigned>* s_cnt,
                                               const ComSlots& slots, Arrive arrive, int* error) {
  const int nm1 = n - 1;
  const int total = 2 * n - 1;
  if ((unsigned)child >= (unsigned)total) {  // only from a corrupt segment
    *error = 1;
    return;
  }
  NodeData cur = load_data(pos_sorted, nm1, nodes, box_lo, box_hi, child);
  int2 u = up[child];
  bool in_global = true;  // cur is also in global memory and visible (an earlier launch wrote it)
  for (int level = 0;; ++level) {
    const int st = climb_status(u, nm1, total, max_pass);
    if (st == 0) break;
    if (st == 1 || level > kMaxClimb) {
      *error = 1;
      break;
    }
    if (st == 2) {  // frontier entry of the next pass (none in the last pass)
      if (seg) push_entry(seg, s_cnt, cur, make_link(up_parent(u), child, u.y), child, error);
      else *error = 1;
      break;
    }
    const int node = up_parent(u), sib = u.y;
    const int2 next = up[node];
    const unsigned slot = __hip_atomic_fetch_add(slots.used, 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_WORKGROUP);
    unsigned tag;
    if (slot < slots.cap) {
      slots.com[slot] = cur.com;
      slots.lo[slot] = make_float3(cur.lo.x, cur.lo.y, cur.lo.z);
      slots.hi[slot] = make_float3(cur.hi.x, cur.hi.y, cur.hi.z);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      tag = slot + 1;
    } else {
      // table full: the sibling reads this child from global memory, so its
      // node stores (if this launch made them) must have completed
      if (!in_global) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      tag = kTagGlobal;
    }
    const unsigned prev = arrive(node, tag);
    if (prev == 0) break;  // first arrival: the sibling combines
    NodeData other;
    if (prev == kTagGlobal) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      other = load_data(pos_sorted, nm1, nodes, box_lo, box_hi, sib);
    } else if (prev - 1 < slots.cap) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const float3 l = slots.lo[prev - 1], h = slots.hi[prev - 1];
      other = NodeData{slots.com[prev - 1], make_float4(l.x, l.y, l.z, 0.f),
                       make_float4(h.x, h.y, h.z, 0.f)};
    } else {  // more than two arrivals: a malformed tree
      *error = 1;
      break;
    }
    cur = combine(cur, other);
    store_node(nodes, box_lo, box_hi, node, cur);
    in_global = false;
    child = node;
    u = next;
  }
}

// Staged climb of passes 2 and 3: every node of the pass is staged in LDS
// with its id and a link to its parent (the parent's staged index and this
// node's side there).  Node k has two child slots; a child arriving at k
// writes its data into its own slot (side 0 for the smaller id of the two
// children), then counts its arrival.  The climb runs in two phases:
//  * arrive_entry: every frontier entry of the previous pass arrives at its
//    staged parent, all threads in parallel; a second arrival puts the node
//    on the ready list;
//  * climb_ready4 (after a barrier): a few waves, one per SIMD, take ready
//    nodes from the list, combine the two slots, store the node and arrive
//    at its parent; a second arrival climbs on, a first one takes the next
//    ready node.
// A climb step is VALU-heavy (the FP64 combine) and each wave issues it for
// all its lanes, so the serial part of the pass costs per level the issue of
// the waves that run it: few full waves beat many sparse ones.
struct Staged {
  lds<float4>* com;  // 2 per node
  lds<float3>* lo;
  lds<float3>* hi;
  lds<int>* id;      // per node
  lds<int>* link;
  lds<unsigned>* arrived;
  lds<int>* ready;   // nodes whose two children have arrived
  lds<unsigned>* nready;
};
// Staged link: parent index | side << 30, with two special parents: kLinkRoot
// (the node is the root) and kOutBase + parent id (the parent belongs to a
// later pass; ids are below 2^28 as n <= 2^28).
constexpr int kSideBit = 1 << kSideShift;
constexpr int kLinkRoot = 0x3fffffff;
constexpr int kOutBase = 1 << 29;
constexpr int kClimbWaves = 4;  // waves of climb_ready4: one per SIMD

// Links of the staged nodes: on entry link[k] holds the node's link from its
// record (parent id | side << 30, -1 at the root); `index(id)` must be ready.
template <typename Index>
__device__ __forceinline__ void link_staged(const Staged& st, int count, Index index) {
  for (int k = threadIdx.x; k < count; k += blockDim.x) {
    const int link = st.link[k];
    int pk = kLinkRoot;
    if (link >= 0) {
      const int parent = link & kIdMask;
      const int ix = index(parent);
      pk = ix >= 0 ? ix : kOutBase + parent;
    }
    st.link[k] = pk | (link >= 0 ? link & kSideBit : 0);
    st.arrived[k] = 0;
  }
}

// `cur` arrives at staged node k on `side`; returns true for the second
// arrival (the node is then complete in LDS).
__device__ __forceinline__ bool arrive_staged(const Staged& st, const NodeData& cur, int k,
                                              int side, int* error) {
  const int mine = 2 * k + side;
  st.com[mine] = cur.com;
  st.lo[mine] = make_float3(cur.lo.x, cur.lo.y, cur.lo.z);
  st.hi[mine] = make_float3(cur.hi.x, cur.hi.y, cur.hi.z);
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  const unsigned prev = __hip_atomic_fetch_add(&st.arrived[k], 1u, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_WORKGROUP);
  if (prev > 1u) *error = 1;  // a third arrival: a malformed tree
  return prev == 1u;
}

// Phase 1 for a frontier entry of the previous pass.  An entry whose parent
// is not staged here belongs to a later pass and is passed on to `out`.
template <typename Index>
__device__ __forceinline__ void arrive_entry(const FrontEntry& e, const Staged& st, Index index,
                                             FrontEntry* out, lds<unsigned>* s_cnt,
                                             int* error) {
  const int link = __float_as_int(e.lo.w);
  const NodeData d{e.com, make_float4(e.lo.x, e.lo.y, e.lo.z, 0.f),
                   make_float4(e.hi.x, e.hi.y, e.hi.z, 0.f)};
  if (link < 0) {  // the root is never an entry
    *error = 1;
    return;
  }
  const int k = index(link & kIdMask);
  if (k < 0) {
    if (out) push_entry(out, s_cnt, d, link, __float_as_int(e.hi.w), error);
    else *error = 1;
    return;
  }
  if (arrive_staged(st, d, k, (link >> kSideShift) & 1, error)) {
    const unsigned r = __hip_atomic_fetch_add(st.nready, 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
    st.ready[r] = k;  // r < count: each node is completed once
  }
}

// Phase 2 with four lanes per climb: lane c of a quad computes component c of
// the node's COM (c = 3: the mass) and AABB, with the operations combine()
// applies to that component, so a climb step issues about a third of the
// instructions of a one-lane step (a lone wave issues one instruction per
// four cycles: the step's instruction count is its latency).  Every decision
// is uniform in the quad: lane 0 makes the arrival count after the whole wave
// has completed its slot writes, and a quad permutation broadcasts it.  Quad
// q takes ready nodes q, q + Q, q + 2Q, ... (Q climbing quads) as its climbs
// end; `count` staged nodes bound the steps of one quad.
template <int kPerm>  // DPP quad_perm control: 0x00 lane 0, 0x55 lane 1, 0xaa lane 2
__device__ __forceinline__ float quad_bcast(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), kPerm, 0xf, 0xf, false));
}
__device__ __forceinline__ void climb_ready4(Node* nodes, float4* box_lo, float4* box_hi,
                                             const Staged& st, int count, FrontEntry* out,
                                             lds<unsigned>* s_cnt, int* error) {
  constexpr unsigned kQuads = kClimbWaves * 16;
  const unsigned nready = *st.nready;
  const int c = threadIdx.x & 3;
  const int cc = c < 3 ? c : 2;  // the AABB component read (lane 3 repeats z)
  unsigned r = threadIdx.x >> 2;
  if (r >= nready) return;  // whole quads
  int k = st.ready[r];
  lds<float>* com = reinterpret_cast<lds<float>*>(st.com);
  lds<float>* lo = reinterpret_cast<lds<float>*>(st.lo);
  lds<float>* hi = reinterpret_cast<lds<float>*>(st.hi);
  // Rare cases (non-finite operands, the root, a parent of a later pass, a
  // malformed tree) are handled under wave-uniform branches, so the common
  // step has no divergent control flow besides the end of a climb.
  for (int step = 0;; ++step) {
    if (step > count) {  // more combines than nodes: a malformed tree
      if (c == 0) *error = 1;
      break;
    }
    int link = st.link[k];
    int node = st.id[k];
    float x0 = com[8 * k + c], x1 = com[8 * k + 4 + c];
    float m0 = com[8 * k + 3], m1 = com[8 * k + 7];
    float l0 = lo[6 * k + cc], l1 = lo[6 * k + 3 + cc];
    float h0 = hi[6 * k + cc], h1 = hi[6 * k + 3 + cc];
    int nk = st.ready[min(r + kQuads, nready - 1)];  // the next ready node, if this climb ends
    // all reads in flight together: one LDS round trip
    asm volatile("" : "+v"(link), "+v"(node), "+v"(x0), "+v"(x1), "+v"(m0), "+v"(m1), "+v"(l0),
                 "+v"(l1), "+v"(h0), "+v"(h1), "+v"(nk));
    const double ma = (double)m0, mb = (double)m1, m = ma + mb;
    const double num = (double)x0 * ma + (double)x1 * mb;
    const bool ok = nr_div_ok(num, m);
    double quot = nr_quot(num, m, nr_recip(m));
    if (__builtin_amdgcn_ballot_w64(!ok)) quot = ok ? quot : num / m;
    const float q = c == 3 ? (float)m : (float)quot;
    const float lmin = fminf(l0, l1), hmax = fmaxf(h0, h1);
    const float edge = hmax - lmin;
    const float sz = fmaxf(quad_bcast<0x00>(edge),
                           fmaxf(quad_bcast<0x55>(edge), quad_bcast<0xaa>(edge)));
    reinterpret_cast<float*>(&nodes[node].com)[c] = q;
    reinterpret_cast<float*>(box_lo + node)[c] = c == 3 ? 0.f : lmin;
    reinterpret_cast<float*>(box_hi + node)[c] = c == 3 ? 0.f : hmax;
    nodes[node].size2 = sz * sz;  // the same value from the four lanes
    const int pk = link & (kSideBit - 1);
    const bool inner = pk < kOutBase;  // the parent is staged in this pass
    int prev = 0;
    if (inner) {
      const int mine = 2 * pk + ((link & kSideBit) ? 1 : 0);
      com[4 * mine + c] = q;
      lo[3 * mine + cc] = lmin;  // lane 3 writes lane 2's value to the same word
      hi[3 * mine + cc] = hmax;
      // the wave's slot writes complete before lane 0 counts the arrival
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      prev = (int)__hip_atomic_fetch_add(&st.arrived[pk], c == 0 ? 1u : 0u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
      prev = __builtin_amdgcn_mov_dpp(prev, 0x00, 0xf, 0xf, false);  // lane 0's count
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }
    if (__builtin_amdgcn_ballot_w64(!inner || prev > 1)) {
      if (inner && prev > 1) {  // a third arrival: a malformed tree
        if (c == 0) *error = 1;
      } else if (!inner && pk != kLinkRoot) {  // the parent belongs to a later pass
        int slot = 0;
        if (c == 0)
          slot = (int)__hip_atomic_fetch_add(s_cnt, 1u, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_WORKGROUP);
        slot = __builtin_amdgcn_mov_dpp(slot, 0x00, 0xf, 0xf, false);
        if (!out || slot >= kSeg) {
          if (c == 0) *error = 1;
        } else {  // the entry's twelve words, three per lane
          const int elink = (pk - kOutBase) | (link & kSideBit);
          float* e = reinterpret_cast<float*>(out + slot);
          e[c] = q;
          e[4 + c] = c == 3 ? __int_as_float(elink) : lmin;
          e[8 + c] = c == 3 ? __int_as_float(node) : hmax;
        }
      }
    }
    if (inner && prev == 1) {
      k = pk;
    } else {  // this climb ended: the quad's next ready node, if any
      r += kQuads;
      if (r >= nready) break;
      k = nk;
    }
  }
}

// Pass-3 node list, filled by pass 2 (one reservation per workgroup) and
// read and reset by pass 3: (id, link) per node.
struct TopList {
  unsigned* count;
  int2* list;
};

// Pass 2: workgroup s owns the superblock of leaves [s*S, s*S + S) and its
// kPer = S / L pass-1 workgroups.  Thread t serves pass-1 workgroup t / 32,
// slots t % 32 (then t % 32 + 32, ... in the rare longer segments): it loads
// that workgroup's two counts, its node record and its frontier entry at the
// start, so the pass waits for one global round trip before it stages.  The
// records of pass 2 are staged (at most kMidCap; usually about 350) with a
// 16-bit id -> index table, those of pass 3 are listed for pass 3.  A
// superblock with more pass-2 nodes climbs with 16-bit arrival words in LDS
// and exchange slots instead (climb_exchange from the entries' ids, reading
// their data from global memory).
constexpr int kMidCap = 640;
constexpr int kMidSlots = 1024;
constexpr size_t kSlotBytes = 16 + 12 + 12;  // COM float4, lo float3, hi float3
constexpr size_t kMidStagedBytes = 2 * kMidCap * kSlotBytes + 4 * kMidCap * 4 + kComSuper * 2;
constexpr size_t kMidUnstagedBytes = kMidSlots * kSlotBytes + kComSuper / 2 * 4;
constexpr size_t kMidBytes = kMidStagedBytes > kMidUnstagedBytes ? kMidStagedBytes
                                                                 : kMidUnstagedBytes;
constexpr int kMidTops = 256;  // pass-3 nodes a superblock collects before listing them
__global__ __launch_bounds__(kComThreads) void com_mid_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int2* __restrict__ up,
    float4* box_lo, float4* box_hi, const FrontEntry* __restrict__ seg1,
    const unsigned* __restrict__ cnt1, const int2* __restrict__ rec1,
    const unsigned* __restrict__ rcnt1, int nblocks1, FrontEntry* seg2, unsigned* cnt2,
    TopList top, int* error) {
  __shared__ __align__(16) unsigned char s_raw[kMidBytes];
  __shared__ int2 s_tops[kMidTops];
  __shared__ unsigned s_nfront, s_used, s_nmid, s_ntop, s_topbase, s_nready;
  const int lo = blockIdx.x * kComSuper;
  const int hi = min(n - 1, lo + kComSuper - 1);
  constexpr int kPer = kComSuper / kComLeaves;  // pass-1 workgroups per superblock
  static_assert(kPer * 32 == kComThreads, "32 threads per pass-1 workgroup");
  // this thread's pass-1 workgroup and first slot; every load is issued
  // before the first is used (clamped, so none waits for a branch)
  const int b1 = blockIdx.x * kPer + (int)(threadIdx.x >> 5);
  const int slot0 = threadIdx.x & 31;
  const int bq = min(b1, nblocks1 - 1);
  const unsigned nrec_raw = rcnt1[bq], nent_raw = cnt1[bq];
  const int2 rec0 = rec1[(size_t)bq * kSeg + slot0];
  const FrontEntry ent0 = seg1[(size_t)bq * kSeg + slot0];
  const unsigned nrec = b1 < nblocks1 ? min(nrec_raw, (unsigned)kSeg) : 0u;
  const unsigned nent = b1 < nblocks1 ? min(nent_raw, (unsigned)kSeg) : 0u;
  // staged layout
  lds<unsigned char>* raw = (lds<unsigned char>*)s_raw;
  Staged st;
  st.com = reinterpret_cast<lds<float4>*>(raw);
  st.lo = reinterpret_cast<lds<float3>*>(st.com + 2 * kMidCap);
  st.hi = st.lo + 2 * kMidCap;
  st.id = reinterpret_cast<lds<int>*>(st.hi + 2 * kMidCap);
  st.link = st.id + kMidCap;
  st.arrived = reinterpret_cast<lds<unsigned>*>(st.link + kMidCap);
  st.ready = reinterpret_cast<lds<int>*>(st.arrived + kMidCap);
  st.nready = (lds<unsigned>*)&s_nready;
  lds<unsigned short>* s_index = reinterpret_cast<lds<unsigned short>*>(st.ready + kMidCap);
  // unstaged layout
  const ComSlots slots{reinterpret_cast<lds<float4>*>(raw),
                       reinterpret_cast<lds<float3>*>(raw + kMidSlots * 16),
                       reinterpret_cast<lds<float3>*>(raw + kMidSlots * 28),
                       (lds<unsigned>*)&s_used, (unsigned)kMidSlots};
  lds<unsigned>* s_arrive = reinterpret_cast<lds<unsigned>*>(raw + kMidSlots * kSlotBytes);

  for (int k = threadIdx.x; k < kComSuper / 2; k += blockDim.x)
    reinterpret_cast<lds<unsigned>*>(s_index)[k] = 0;
  if (threadIdx.x == 0) {
    s_nfront = 0;
    s_used = 0;
    s_nmid = 0;
    s_ntop = 0;
    s_nready = 0;
  }
  __syncthreads();
  for (unsigned j = slot0; j < nrec; j += 32) {
    const int2 r = j == (unsigned)slot0 ? rec0 : rec1[(size_t)bq * kSeg + j];
    const unsigned pass = (unsigned)r.x >> 30;
    const int id = r.x & kIdMask;
    if (pass == 1 && id >= lo && id <= hi) {
      const unsigned k = __hip_atomic_fetch_add(&s_nmid, 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
      if (k < (unsigned)kMidCap) {
        st.id[k] = id;
        st.link[k] = r.y;
        s_index[id - lo] = (unsigned short)(k + 1);
      }
    } else if (pass == 2) {
      const unsigned k = __hip_atomic_fetch_add(&s_ntop, 1u, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_WORKGROUP);
      if (k < (unsigned)kMidTops) {
        s_tops[k] = make_int2(id, r.y);
      } else {  // many pass-3 nodes in one superblock: list this one alone
        const unsigned g = atomic_add_relaxed(top.count, 1u);
        if (g < (unsigned)kTopList) top.list[g] = make_int2(id, r.y);
      }
    } else {  // a record of pass 1, or of a node outside the superblock
      *error = 1;
    }
  }
  __syncthreads();
  const unsigned nmid = s_nmid;
  const bool staged = nmid <= (unsigned)kMidCap;
  const unsigned ntop = min(s_ntop, (unsigned)kMidTops);
  if (threadIdx.x == 0 && ntop > 0) s_topbase = atomic_add_relaxed(top.count, ntop);
  auto index = [&](int id) {
    const int rel = id - lo;
    return (rel >= 0 && rel < kComSuper) ? (int)s_index[rel] - 1 : -1;
  };
  if (staged) {
    link_staged(st, (int)nmid, index);
  } else {
    for (int k = threadIdx.x; k < kComSuper / 2; k += blockDim.x) s_arrive[k] = 0;
  }
  __syncthreads();
  if (ntop > 0 && threadIdx.x < ntop && s_topbase + threadIdx.x < (unsigned)kTopList)
    top.list[s_topbase + threadIdx.x] = s_tops[threadIdx.x];
  auto arrive = [&](int node, unsigned tag) {
    const int k = node - lo;
    const unsigned shift = 16u * (unsigned)(k & 1);
    const unsigned old = __hip_atomic_fetch_or(&s_arrive[k >> 1], tag << shift,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    return (old >> shift) & 0xffffu;
  };
  FrontEntry* out = seg2 + (size_t)blockIdx.x * kSeg;
  lds<unsigned>* nfront = (lds<unsigned>*)&s_nfront;
  for (unsigned j = slot0; j < nent; j += 32) {
    const FrontEntry e = j == (unsigned)slot0 ? ent0 : seg1[(size_t)bq * kSeg + j];
    if (staged)
      arrive_entry(e, st, index, out, nfront, error);
    else
      climb_exchange(pos_sorted, n, nodes, up, box_lo, box_hi, __float_as_int(e.hi.w), 1u, out,
                     nfront, slots, arrive, error);
  }
  if (staged) {  // uniform: nmid is one value for the block
    __syncthreads();
    if (threadIdx.x < kClimbWaves * 64)
      climb_ready4(nodes, box_lo, box_hi, st, (int)nmid, out, nfront, error);
  }
  __syncthreads();
  if (threadIdx.x == 0) cnt2[blockIdx.x] = min(s_nfront, (unsigned)kSeg);
}

// Pass 3: one workgroup climbs from every superblock's frontier segment
// through the nodes that span superblocks (about 520 at 1M leaves), staged
// from the list pass 2 made, with an LDS hash from id to index.  Thread t
// serves segment t / 32 (+ 32, 64, ...), slots t % 32 (+ 32, ...), and its
// first entry and list record are loaded at the start.  With more than
// kTopCap listed nodes it climbs with arrival tags in the low bits of the
// per-node `visit` words in global memory (written by the Karras kernel with
// those bits clear; each node's word is touched by one workgroup only, so
// workgroup scope holds) and exchange slots instead.  Earlier passes are
// visible across the kernel boundaries.
//
// A forest's systems are independent trees (the nodes above the system roots
// get no COM), so its pass 3 runs as kTopGroups workgroups: workgroup w takes
// systems [w S / W, (w + 1) S / W), i.e. the listed nodes whose ids lie in
// those systems' leaf range (an internal node's id is one end of its leaf
// range) and the entries whose parents do, from the superblocks that range
// overlaps.  Each stages its own share (the 256-system ensemble lists ~2900
// nodes, beyond one workgroup's kTopCap).  Pass 1 resets the list count for
// the next build, when every workgroup here has read it.
constexpr int kTopGroups = 16;
constexpr int kTopCap = 1024;
constexpr int kTopHash = 2048;  // power of two
constexpr int kTopSlots = 2048;
constexpr size_t kTopStagedBytes = 2 * kTopCap * kSlotBytes + 4 * kTopCap * 4 + 2 * kTopHash * 4;
constexpr size_t kTopUnstagedBytes = kTopSlots * kSlotBytes;
constexpr size_t kTopBytes = kTopStagedBytes > kTopUnstagedBytes ? kTopStagedBytes
                                                                 : kTopUnstagedBytes;
__device__ __forceinline__ unsigned top_hash(int id) {
  return ((unsigned)id * 2654435761u) >> (32 - 11);
}
static_assert(kTopHash == 1 << 11, "top_hash yields 11 bits");
static_assert(kTopCap <= kComThreads, "one listed node per thread in the first load");

__global__ __launch_bounds__(kComThreads) void com_top_kernel(
    const float4* __restrict__ pos_sorted, int n, Node* nodes, const int2* __restrict__ up,
    unsigned* visit, float4* box_lo, float4* box_hi, const FrontEntry* __restrict__ seg2,
    const unsigned* __restrict__ cnt2, int nblocks2, TopList top,
    const int* __restrict__ sys_off, int nsys, int* error) {
  __shared__ __align__(16) unsigned char s_raw[kTopBytes];
  __shared__ unsigned s_used, s_nready, s_nown;
  const bool part = gridDim.x > 1;  // a forest split by systems (uniform)
  // node ids [own_lo, own_hi) and superblocks [bfirst, blast] of this workgroup
  int own_lo = 0, own_hi = n, bfirst = 0, blast = nblocks2 - 1;
  if (part) {
    const int s0 = (int)((long long)blockIdx.x * nsys / gridDim.x);
    const int s1 = (int)((long long)(blockIdx.x + 1) * nsys / gridDim.x);
    own_lo = sys_off[s0];
    own_hi = sys_off[s1];
    bfirst = own_lo / kComSuper;
    blast = min(nblocks2 - 1, max(own_hi - 1, 0) / kComSuper);
  }
  auto owns = [&](int id) { return id >= own_lo && id < own_hi; };
  const unsigned listed_all = *top.count;
  // staging needs the whole list, and at most kTopCap nodes (known for one
  // workgroup now, for a share of a forest after the selection below)
  const bool may_stage = listed_all <= (unsigned)(part ? kTopList : kTopCap);
  // first loads: this thread's listed node and its first entry
  const int2 rec0 = top.list[threadIdx.x];
  const int b0 = bfirst + (int)(threadIdx.x >> 5), slot0 = threadIdx.x & 31;
  const int bq = max(0, min(b0, blast));
  const unsigned nent0_raw = cnt2[bq];
  const FrontEntry ent0 = seg2[(size_t)bq * kSeg + slot0];
  lds<unsigned char>* raw = (lds<unsigned char>*)s_raw;
  Staged st;
  st.com = reinterpret_cast<lds<float4>*>(raw);
  st.lo = reinterpret_cast<lds<float3>*>(st.com + 2 * kTopCap);
  st.hi = st.lo + 2 * kTopCap;
  st.id = reinterpret_cast<lds<int>*>(st.hi + 2 * kTopCap);
  st.link = st.id + kTopCap;
  st.arrived = reinterpret_cast<lds<unsigned>*>(st.link + kTopCap);
  st.ready = reinterpret_cast<lds<int>*>(st.arrived + kTopCap);
  st.nready = (lds<unsigned>*)&s_nready;
  lds<int>* s_key = reinterpret_cast<lds<int>*>(st.ready + kTopCap);
  lds<int>* s_val = s_key + kTopHash;
  const ComSlots slots{reinterpret_cast<lds<float4>*>(raw),
                       reinterpret_cast<lds<float3>*>(raw + kTopSlots * 16),
                       reinterpret_cast<lds<float3>*>(raw + kTopSlots * 28),
                       (lds<unsigned>*)&s_used, (unsigned)kTopSlots};
  if (threadIdx.x == 0) {
    s_used = 0;
    s_nready = 0;
    s_nown = 0;
  }
  if (may_stage)
    for (int k = threadIdx.x; k < kTopHash; k += blockDim.x) s_key[k] = -1;
  __syncthreads();
  unsigned listed = listed_all;
  int2 mine = rec0;  // the listed node thread k stages at index k
  if (part) {
    if (may_stage)
      for (unsigned k = threadIdx.x; k < listed_all; k += blockDim.x) {
        const int2 r = k == threadIdx.x ? rec0 : top.list[k];
        if (!owns(r.x)) continue;
        const unsigned ix = __hip_atomic_fetch_add(&s_nown, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
        if (ix < (unsigned)kTopCap) {
          st.id[ix] = r.x;
          st.link[ix] = r.y;
        }
      }
    __syncthreads();
    listed = s_nown;
    if (threadIdx.x < min(listed, (unsigned)kTopCap))
      mine = make_int2(st.id[threadIdx.x], st.link[threadIdx.x]);
  }
  const bool staged = may_stage && listed <= (unsigned)kTopCap;
  if (staged && threadIdx.x < listed) {
    const int k = threadIdx.x;
    const int id = mine.x;
    if (id < 0 || id >= n - 1) {  // only on a stale or corrupt list
      *error = 1;
      st.id[k] = 0;  // keeps every later access in range
      st.link[k] = -1;
    } else {
      st.id[k] = id;
      st.link[k] = mine.y;
      unsigned h = top_hash(id);
      bool placed = false;
      for (int probe = 0; probe < kTopHash && !placed; ++probe) {
        const int old = atomicCAS(&s_key[h], -1, id);
        if (old == -1) {
          s_val[h] = k;
          placed = true;
        } else if (old == id) {  // listed twice: a stale or corrupt list
          *error = 1;
          placed = true;
        }
        h = (h + 1) & (kTopHash - 1);
      }
    }
  }
  __syncthreads();
  auto index = [&](int id) {
    unsigned h = top_hash(id);
    for (int probe = 0; probe < kTopHash; ++probe) {
      const int key = s_key[h];
      if (key == id) return s_val[h];
      if (key == -1) break;
      h = (h + 1) & (kTopHash - 1);
    }
    return -1;
  };
  if (staged) {
    link_staged(st, (int)listed, index);
    __syncthreads();
  }
  auto arrive = [&](int node, unsigned tag) {
    return __hip_atomic_fetch_or(&visit[node], tag, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP) & 0xffffu;
  };
  // the whole tree (or this workgroup's systems) is in range here: an entry
  // or climb whose parent is not listed means a corrupt list
  for (int b = b0; b <= blast; b += kComThreads / 32) {
    const unsigned nent = min(b == b0 ? nent0_raw : cnt2[b], (unsigned)kSeg);
    for (unsigned j = slot0; j < nent; j += 32) {
      const FrontEntry e =
          (b == b0 && j == (unsigned)slot0) ? ent0 : seg2[(size_t)b * kSeg + j];
      const int elink = __float_as_int(e.lo.w);
      if (part && elink >= 0 && !owns(elink & kIdMask)) continue;  // another share's
      if (staged)
        arrive_entry(e, st, index, nullptr, nullptr, error);
      else
        climb_exchange(pos_sorted, n, nodes, up, box_lo, box_hi, __float_as_int(e.hi.w), 2u,
                       nullptr, nullptr, slots, arrive, error);
    }
  }
  if (staged) {  // uniform: listed is one value for the block
    __syncthreads();
    if (threadIdx.x < kClimbWaves * 64)
      climb_ready4(nodes, box_lo, box_hi, st, (int)listed, nullptr, nullptr, error);
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
  const bool big = n >= (1 << 19);
  const int g = grid_for(n, big ? 16 : 8, 2048);
  if (key_bits == 30) {
    if (big)
      morton_hist_kernel<uint32_t, 16><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                            (uint32_t*)keys, ss);
    else
      morton_hist_kernel<uint32_t, 8><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                           (uint32_t*)keys, ss);
  } else {
    if (big)
      morton_hist_kernel<uint64_t, 16><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                            (uint64_t*)keys, ss);
    else
      morton_hist_kernel<uint64_t, 8><<<g, kBlock, 0, s>>>(pos, n, partial, nparts, qp,
                                                           (uint64_t*)keys, ss);
  }
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

void launch_karras(hipStream_t s, int key_bits, const void* keys, int n, Node* nodes, int2* up,
                   unsigned* visit, SortState* ss) {
  if (n < 2) return;
  int g = (n - 1 + kBlock - 1) / kBlock;
  if (key_bits == 30)
    karras_kernel<uint32_t><<<g, kBlock, 0, s>>>((const uint32_t*)keys, n, nodes, up, visit, ss);
  else
    karras_kernel<uint64_t><<<g, kBlock, 0, s>>>((const uint64_t*)keys, n, nodes, up, visit, ss);
  NB_LAUNCH_CHECK();
}

namespace {
int com_g1(int n) { return (n + kComLeaves - 1) / kComLeaves; }
int com_g2(int n) { return (n + kComSuper - 1) / kComSuper; }
}  // namespace

int com_scratch_ints(int n) {
  const size_t g1 = (size_t)com_g1(n), g2 = (size_t)com_g2(n);
  const size_t ints = kComCounterInts + 2 * (size_t)kTopList +
                      (g1 + g2) * kSeg * (sizeof(FrontEntry) / 4) + g1 * kSeg * 2 + 2 * g1 + g2;
  return (int)ints;
}

void launch_com(hipStream_t s, float4* pos_sorted, int n, Node* nodes, const int2* up,
                unsigned* visit, float4* box_lo, float4* box_hi, int* scratch, int* error,
                const unsigned* perm, const float4* gather_src, const int* sys_off,
                int nsys) {
  if (n < 2) {  // no tree; a single particle is its own sorted position
    if (n == 1 && gather_src)
      NB_HIP(hipMemcpyAsync(pos_sorted, gather_src, sizeof(float4), hipMemcpyDeviceToDevice, s));
    return;
  }
  const int g1 = com_g1(n), g2 = com_g2(n);
  if (g2 > kTopSegs) throw Error(kErrInvalid, "particle count too large for the COM top pass");
  static_assert(kComCounterInts % 4 == 0 && sizeof(FrontEntry) == 48, "16-byte aligned carve");
  const TopList top{reinterpret_cast<unsigned*>(scratch),
                    reinterpret_cast<int2*>(scratch + kComCounterInts)};
  FrontEntry* seg1 = reinterpret_cast<FrontEntry*>(top.list + kTopList);
  FrontEntry* seg2 = seg1 + (size_t)g1 * kSeg;
  int2* rec1 = reinterpret_cast<int2*>(seg2 + (size_t)g2 * kSeg);
  unsigned* cnt1 = reinterpret_cast<unsigned*>(rec1 + (size_t)g1 * kSeg);
  unsigned* rcnt1 = cnt1 + g1;
  unsigned* cnt2 = rcnt1 + g1;
  com_local_kernel<<<g1, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, visit, box_lo, box_hi,
                                              seg1, cnt1, rec1, rcnt1, perm, gather_src,
                                              top.count, error);
  NB_LAUNCH_CHECK();
  if (g1 == 1) return;
  com_mid_kernel<<<g2, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, box_lo, box_hi, seg1, cnt1,
                                            rec1, rcnt1, g1, seg2, cnt2, top, error);
  NB_LAUNCH_CHECK();
  if (g2 == 1) return;
  const int wtop = sys_off && nsys > 1 ? std::min(nsys, kTopGroups) : 1;
  com_top_kernel<<<wtop, kComThreads, 0, s>>>(pos_sorted, n, nodes, up, visit, box_lo, box_hi,
                                              seg2, cnt2, g2, top, sys_off, nsys, error);
  NB_LAUNCH_CHECK();
}

}  // namespace nb
