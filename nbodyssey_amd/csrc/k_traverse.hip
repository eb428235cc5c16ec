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
//   walk_target_kernel    the per-target walk of traverse_kernel: every target
//                         keeps its own decisions, and a wave walks its 64
//                         targets together with a lane set per stack entry.
//   walk_wave_kernel      the group walk of traverse_warp_kernel at group width
//                         64, 32 or 16, one group per wave64: a node opens if
//                         any live lane of the group opens it; node-parallel
//                         batches (up to 64 nodes, one per lane) settle the
//                         votes, and an LDS interaction list is evaluated by
//                         all lanes (walk_group64 below).  Same visited /
//                         opened / accepted sets, hence the same counts, as the
//                         one-node-at-a-time warp vote.
//   walk_split_kernel     group widths 32 and 16 in FP32 from 2^19 targets on:
//                         64 targets per wave in 2 or 4 subgroups that share
//                         the node batches but keep their own votes and
//                         interaction lists.
//   walk_items_kernel     the wave64 group walk over ensemble work items pulled from a
//                         device queue (persistent).
//   walk_potential_kernel the same walk accumulating the potential (tree energy).
// COUNT variants also record interactions and opening tests (per target for
// the thread walk, per group otherwise).
#include "nb_device.hpp"
#include "nb_internal.hpp"

#include <algorithm>

namespace nb {
namespace {

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Renumber them so that each XCD runs runs of kXcdChunk consecutive
// workgroups (Morton-adjacent targets, dealt to the XCDs in turn):
// neighbouring groups open nearly the same nodes, so an XCD's L2 serves a run
// from one region's part of the tree.  Runs of 4-64 workgroups measure the
// same; whole eighths of the targets per XCD lose to the imbalance between
// the XCDs' shares of the work.
constexpr int kXcds = 8;
constexpr int kXcdChunk = 16;  // consecutive workgroups kept on one XCD
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, g = gridDim.x;
  const int span = kXcds * kXcdChunk;
  if (b >= (g / span) * span) return b;
  const int x = b % kXcds, k = b / kXcds;
  return ((k / kXcdChunk) * kXcds + x) * kXcdChunk + k % kXcdChunk;
}

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

// Node-parallel wave64 group walk.  The visited / opened / accepted node sets
// are exactly those of the one-node-at-a-time group vote (a node is visited
// iff all its ancestors were opened, and each open decision depends only on
// the node and the group), so the interaction and opening-test counts are
// identical; only the FP32 summation order differs.
//
// Per iteration the wave pops up to 64 node ids from its LDS stack, lane k
// loads node k (independent vector loads), and the open/accept vote of each
// node over the 64 targets is settled by
//   * a prefilter on the group's bounding box: a node is accepted by every
//     lane if size2 < theta2 * dmin2 (dmin = COM-to-box distance, with a
//     1e-5 relative margin so FP32 rounding cannot flip a lane's test);
//   * representative targets (below): one whose own canonical test opens the
//     node opens it for the group.  A farthest-corner test ("every lane
//     opens") adds nothing here: every representative lies in the box, so
//     whenever it holds a representative opens the node too;
//   * otherwise the exact per-lane test in the canonical order, node by node
//     with the node broadcast from LDS, reduced by ballot.
// Opened nodes push their two children; accepted nodes and leaves append
// their (x, y, z, m) to an LDS interaction list that all 64 lanes evaluate in
// blocks of 64 (LDS broadcast reads, pure FP32 VALU), optionally Kahan-adding
// each block's partial sums (compensated mode).
constexpr int kWaveStack = 1024;   // LDS stack entries per wave (default)
// Per-target walk: 768 stack entries (3 KB, plus 3 KB of lane sets) let four
// workgroups share a CU instead of three.  Its stack peaks near 260 entries at
// theta 0.5 and 500 at theta 0.3 on 1M particles (1M step: 1.90 -> 1.74 ms at
// theta 0.5, 3.94 -> 3.90 ms at 0.3); 640 entries push theta 0.3 into the
// pair-at-a-time mode near the top (6.0 ms).
constexpr int kTargetStack = 768;
constexpr int kStackReserve = 64 + 2 * kStackDepth;
constexpr int kWaveList = 128;     // LDS interaction-list entries per wave
constexpr float kMacMargin = 1e-5f;
// Representative targets for the open shortcut: 8 per wave for the split
// walk (4 or 2 per subgroup), 6 for the 64-wide group walk (1M two-disk,
// theta 0.5: 6 -> 0.626 ms per walk, 8 -> 0.636 ms; each one costs six VALU
// per batch, and the exact vote it saves costs about eight per node).
constexpr int kReps = 8;
constexpr int kGroupReps = 6;

// POT: also accumulate the potential sum of m / sqrt(r^2 + eps^2) into pot
// (m * inv is the first product of the force factor, so the force bits are
// unchanged).
template <bool COMP, bool POT>
__device__ __forceinline__ void flush_list(const float4* list, int cnt, float bx, float by,
                                           float bz, float eps2, float& ax, float& ay, float& az,
                                           float& cx, float& cy, float& cz, float& pot) {
  float px = 0.f, py = 0.f, pz = 0.f;
  if (!COMP) { px = ax; py = ay; pz = az; }
  if (POT) {
#pragma unroll 8
    for (int k = 0; k < cnt; ++k) {
      const float4 src = list[k];
      const float dx = src.x - bx, dy = src.y - by, dz = src.z - bz;
      const float inv = rsq(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2))));
      const float mi = src.w * inv;
      const float f = mi * inv * inv;
      px = fmaf(dx, f, px);
      py = fmaf(dy, f, py);
      pz = fmaf(dz, f, pz);
      pot += mi;
    }
  } else {
#pragma unroll 8
    for (int k = 0; k < cnt; ++k) accumulate_src(list[k], bx, by, bz, eps2, px, py, pz);
  }
  if (COMP) {
    kahan_add(ax, cx, px);
    kahan_add(ay, cy, py);
    kahan_add(az, cz, pz);
  } else {
    ax = px; ay = py; az = pz;
  }
}

// FP16 far field (precision mode 2).  Accepted cells whose magnitudes are
// safely inside FP16 range are evaluated two per lane-instruction with packed
// FP16 arithmetic, positions relative to the group's box centre and masses
// scaled by 512; every product is accumulated in FP32 (mixed FMA) and scaled
// back once.  A cell goes to this list only if, with h the group box's
// half-extent on each axis,
//   |cell - centre| + h < 128 per axis: every lane's |d| < 128 per axis, so
//       r^2 = dx^2 + dy^2 + dz^2 + eps^2 < 3 * 16384 + eps^2 < 65504;
//   2^-14 / 512 <= m <= 64: m * 512 is a normal FP16 number below 32768;
//   m / (dmin^2 + eps^2) <= 32 (dmin = distance to the group box): every
//       lane has r^2 >= dmin^2 + eps^2, so m * 512 * inv^2 <= 16384.
// Everything else (leaves, near or heavy cells) stays in the FP32 list.
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
constexpr float kHalfMassScale = 512.f;
constexpr float kHalfMaxAccel = 32.f;
constexpr float kHalfMaxSep = 128.f;
constexpr float kHalfMaxMass = 64.f;
constexpr float kHalfMinMass = 6.103515625e-05f / kHalfMassScale;  // 2^-14 / 512
constexpr int kHalfList = 128;   // FP16 list entries per wave (64 pairs)

// l16: pairs of entries, pair p = half2 (x0,x1), (y0,y1), (z0,z1), (m0,m1) at
// l16[4p..4p+3]; npairs pairs.
__device__ __forceinline__ void flush_half(const half2_t* l16, int npairs, half2_t bx, half2_t by,
                                           half2_t bz, half2_t eps2h, float& fx, float& fy,
                                           float& fz) {
#pragma unroll 4
  for (int p = 0; p < npairs; ++p) {
    const half2_t X = l16[4 * p], Y = l16[4 * p + 1], Z = l16[4 * p + 2], M = l16[4 * p + 3];
    const half2_t dx = X - bx, dy = Y - by, dz = Z - bz;
    const half2_t r2 = __builtin_elementwise_fma(
        dz, dz, __builtin_elementwise_fma(dy, dy, __builtin_elementwise_fma(dx, dx, eps2h)));
    half2_t inv;
    inv.x = __builtin_amdgcn_rsqh(r2.x);
    inv.y = __builtin_amdgcn_rsqh(r2.y);
    const half2_t w = M * (inv * inv);
    const half2_t ux = dx * inv, uy = dy * inv, uz = dz * inv;
    fx = fmaf((float)ux.x, (float)w.x, fx);
    fy = fmaf((float)uy.x, (float)w.x, fy);
    fz = fmaf((float)uz.x, (float)w.x, fz);
    fx = fmaf((float)ux.y, (float)w.y, fx);
    fy = fmaf((float)uy.y, (float)w.y, fy);
    fz = fmaf((float)uz.y, (float)w.y, fz);
  }
}

// Per-target walk (traverse_kernel's semantics: every target its own
// depth-first walk and its own open/accept decisions) run by one wave over 64
// consecutive sorted targets.  A stack entry carries the set of lanes for
// which all of the node's ancestors were opened, so the wave loads and tests
// each node once for all of them, in batches of up to 64 nodes, instead of
// every lane chasing its own chain of dependent node fetches.  Per node:
//   * a leaf interacts with every lane of its set;
//   * an internal node is settled for the whole set by the group box when
//     size2 < theta2 * dmin2 (all accept) or size2 > theta2 * dmax2 (all
//     open), each with the 1e-5 margin against FP32 rounding; otherwise every
//     lane runs its own canonical test and the ballot is the set that opens;
//   * the lanes that open it push its two children with that set.
// Children are pushed in pairs that share one set (the root shares its pair
// with a null entry), so a set is stored per pair.  Accepted nodes and leaves
// go to an interaction list with their sets; each lane adds only the entries
// of its own set (the others contribute an exact zero), so every target sums
// exactly its own interaction list, in batch order.
__device__ __forceinline__ void flush_masked(const float4* list, const unsigned* mw, int cnt,
                                             int bit, float bx, float by, float bz, float eps2,
                                             float& ax, float& ay, float& az, int& taken) {
#pragma unroll 8
  for (int k = 0; k < cnt; ++k) {
    const float4 s = list[k];
    // all ones iff this lane is in the set (one signed bit-field extract)
    const int keep = __builtin_amdgcn_sbfe((int)mw[k], (unsigned)bit, 1u);
    const float dx = s.x - bx, dy = s.y - by, dz = s.z - bz;
    const float inv = rsq(fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, eps2))));
    const float f = __int_as_float(__float_as_int(s.w * inv * inv * inv) & keep);
    ax = fmaf(dx, f, ax);
    ay = fmaf(dy, f, ay);
    az = fmaf(dz, f, az);
    taken -= keep;
  }
}

template <bool COUNT>
__device__ __forceinline__ void walk_target64(
    int t0, int count, const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm,
    int n, const Node* __restrict__ nodes, float theta2, float eps2, float G,
    float4* __restrict__ acc, float4* __restrict__ vel, float half_dt, int* __restrict__ counts,
    int* __restrict__ opens, int* overflow, int* error, int* stk, unsigned long long* smask,
    float4* list, unsigned* mlo, unsigned* mhi) {
  const int lane = threadIdx.x & 63;
  const bool live = lane < count;
  const int t = t0 + (live ? lane : count - 1);
  const int nm1 = n - 1;
  const float4 b = pos_sorted[t];
  const unsigned long long live_mask = __builtin_amdgcn_ballot_w64(live);
  const float blx = wave_min(b.x), bly = wave_min(b.y), blz = wave_min(b.z);
  const float bhx = wave_max(b.x), bhy = wave_max(b.y), bhz = wave_max(b.z);
  const float acc_lo = theta2 * (1.f - kMacMargin);
  const float acc_hi = theta2 * (1.f + kMacMargin);
  // this lane's word of a 64-bit set and its bit in it
  const unsigned* mw = lane < 32 ? mlo : mhi;
  const int bit = lane & 31;

  float ax = 0.f, ay = 0.f, az = 0.f;
  int taken = 0, tested = 0;
  int sp = 2, nlist = 0;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  if (lane == 0) {
    stk[0] = 0;   // root
    stk[1] = -1;  // null partner
    smask[0] = live_mask;
  }
  while (sp > 0) {
    // sp stays even: pairs are popped whole (near the top of the stack one
    // pair at a time, two entries per tree level, so it cannot overflow)
    int k = sp < 64 ? sp : 64;
    if (sp > kTargetStack - kStackReserve) k = 2;
    // lane l < k takes stk[sp-1-l] and its pair's set; the other lanes re-read
    // the last popped entry (in bounds) with an empty set
    const bool in_batch = lane < k;
    const int slot = max(sp - 1 - lane, sp - k);
    const int id = stk[slot];
    const unsigned long long m = in_batch ? smask[slot >> 1] : 0ull;
    sp -= k;
    pops += k;  // every node is pushed once, by its parent, with one set
    const bool valid = in_batch && id >= 0;
    if (__builtin_amdgcn_ballot_w64(valid && (unsigned)id >= nodes_total) ||
        pops > nodes_total + 1u) {
      if (lane == 0) *error = 1;
      break;
    }
    const bool is_leaf = valid && id >= nm1;
    const bool is_node = valid && id < nm1;
    // (x, y, z, m) of the leaf or node and the node's (left, right, size2),
    // loaded unconditionally (other lanes read node 0, unused)
    const int nid = is_node ? id : 0;
    const float4* srcp = is_leaf ? pos_sorted + (id - nm1) : &nodes[nid].com;
    const float4 src = *srcp;
    const int4 meta = *reinterpret_cast<const int4*>(&nodes[nid].left);
    const float s2 = __int_as_float(meta.z);
    const int left = meta.x, right = meta.y;
    const float ex = fmaxf(fmaxf(blx - src.x, src.x - bhx), 0.f);
    const float ey = fmaxf(fmaxf(bly - src.y, src.y - bhy), 0.f);
    const float ez = fmaxf(fmaxf(blz - src.z, src.z - bhz), 0.f);
    const float dmin2 = ex * ex + ey * ey + ez * ez;
    const float fx = fmaxf(fabsf(src.x - blx), fabsf(src.x - bhx));
    const float fy = fmaxf(fabsf(src.y - bly), fabsf(src.y - bhy));
    const float fz = fmaxf(fabsf(src.z - blz), fabsf(src.z - bhz));
    const float dmax2 = fx * fx + fy * fy + fz * fz;
    const bool all_acc = s2 < acc_lo * dmin2;
    const bool all_open = s2 > acc_hi * dmax2;
    const bool amb = is_node && !(all_acc || all_open);
    unsigned long long om = is_node && all_open ? m : 0ull;  // lanes of the set that open it
    const unsigned long long ambm = __builtin_amdgcn_ballot_w64(amb);
    if (ambm) {
      // per-lane canonical tests of the unsettled nodes, staged in the free
      // upper half of the list (< 64 entries pending), four per iteration and
      // the rest one at a time; lane 0 stores the q-th node's ballot in the
      // free upper halves of the set words, where the node's lane reads it
      float4* vbuf = list + 64;
      const int j = mbcnt64(ambm);
      if (amb) vbuf[j] = make_float4(src.x, src.y, src.z, s2);
      const int na = __popcll(ambm);
      int q = 0;
      for (; q + 4 <= na; q += 4) {
        const float4 v0 = vbuf[q], v1 = vbuf[q + 1], v2 = vbuf[q + 2], v3 = vbuf[q + 3];
        const unsigned long long b0 = __builtin_amdgcn_ballot_w64(
            !(v0.w < theta2 * dist2(v0.x - b.x, v0.y - b.y, v0.z - b.z)));
        const unsigned long long b1 = __builtin_amdgcn_ballot_w64(
            !(v1.w < theta2 * dist2(v1.x - b.x, v1.y - b.y, v1.z - b.z)));
        const unsigned long long b2 = __builtin_amdgcn_ballot_w64(
            !(v2.w < theta2 * dist2(v2.x - b.x, v2.y - b.y, v2.z - b.z)));
        const unsigned long long b3 = __builtin_amdgcn_ballot_w64(
            !(v3.w < theta2 * dist2(v3.x - b.x, v3.y - b.y, v3.z - b.z)));
        if (lane == 0) {  // q is a multiple of 4: 16-byte aligned
          *reinterpret_cast<uint4*>(mlo + 64 + q) =
              make_uint4((unsigned)b0, (unsigned)b1, (unsigned)b2, (unsigned)b3);
          *reinterpret_cast<uint4*>(mhi + 64 + q) =
              make_uint4((unsigned)(b0 >> 32), (unsigned)(b1 >> 32), (unsigned)(b2 >> 32),
                         (unsigned)(b3 >> 32));
        }
      }
      for (; q < na; ++q) {
        const float4 v = vbuf[q];
        const unsigned long long bq = __builtin_amdgcn_ballot_w64(
            !(v.w < theta2 * dist2(v.x - b.x, v.y - b.y, v.z - b.z)));
        if (lane == 0) {
          mlo[64 + q] = (unsigned)bq;
          mhi[64 + q] = (unsigned)(bq >> 32);
        }
      }
      if (amb) om = (((unsigned long long)mhi[64 + j] << 32) | mlo[64 + j]) & m;
    }
    const unsigned long long am = is_leaf ? m : is_node ? (m & ~om) : 0ull;
    if (COUNT) {
      // opening tests: internal nodes of the batch whose set holds this lane
      const unsigned long long nm = __builtin_amdgcn_ballot_w64(is_node);
      if (is_node) {
        const int r = mbcnt64(nm);
        mlo[64 + r] = (unsigned)m;
        mhi[64 + r] = (unsigned)(m >> 32);
      }
      const int nn = __popcll(nm);
      for (int r = 0; r < nn; ++r) tested += (int)((mw[64 + r] >> bit) & 1u);
    }
    // push the children of nodes some lane opened, as a pair with that set
    const unsigned long long pm = __builtin_amdgcn_ballot_w64(om != 0ull);
    const int npush = __popcll(pm);
    if (sp + 2 * npush > kTargetStack) {
      if (lane == 0) *overflow = 1;
      break;
    }
    if (__builtin_amdgcn_inverse_ballot_w64(pm)) {
      const int ps = sp + 2 * mbcnt64(pm);
      stk[ps] = left;
      stk[ps + 1] = right;
      smask[ps >> 1] = om;
    }
    sp += 2 * npush;
    // append accepted nodes and leaves with their sets (lane order)
    const unsigned long long apm = __builtin_amdgcn_ballot_w64(am != 0ull);
    if (__builtin_amdgcn_inverse_ballot_w64(apm)) {
      const int e = nlist + mbcnt64(apm);
      list[e] = src;
      mlo[e] = (unsigned)am;
      mhi[e] = (unsigned)(am >> 32);
    }
    nlist += __popcll(apm);
    if (nlist >= 64) {
      flush_masked(list, mw, 64, bit, b.x, b.y, b.z, eps2, ax, ay, az, taken);
      nlist -= 64;
      if (lane < nlist) {
        list[lane] = list[64 + lane];
        mlo[lane] = mlo[64 + lane];
        mhi[lane] = mhi[64 + lane];
      }
    }
  }
  if (nlist > 0) flush_masked(list, mw, nlist, bit, b.x, b.y, b.z, eps2, ax, ay, az, taken);
  if (COUNT && live) {
    counts[t] = taken;
    opens[t] = tested;
  }
  if (live) finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

template <bool COUNT>
__global__ __launch_bounds__(kBlock) void walk_target_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error) {
  constexpr int WAVES = kBlock / 64;
  __shared__ int s_stack[WAVES][kTargetStack];
  __shared__ unsigned long long s_smask[WAVES][kTargetStack / 2];
  __shared__ float4 s_list[WAVES][kWaveList];
  __shared__ __align__(16) unsigned s_mlo[WAVES][kWaveList];
  __shared__ __align__(16) unsigned s_mhi[WAVES][kWaveList];
  const int w = threadIdx.x >> 6;
  const int gbase = xcd_block() * blockDim.x + (threadIdx.x & ~63);
  if (gbase >= n) return;  // wave-uniform
  walk_target64<COUNT>(gbase, min(64, n - gbase), pos_sorted, perm, n, nodes, theta2, eps2, G,
                       acc, vel, half_dt, counts, opens, overflow, error, s_stack[w], s_smask[w],
                       s_list[w], s_mlo[w], s_mhi[w]);
}

// One group of `count` (<= 64) consecutive sorted targets starting at t0,
// walking the subtree under `root`, on one wave; `slot` indexes the count
// outputs.  Dead lanes (lane >= count) repeat the group's last target, so the
// group box and the representatives are those of the live targets.
// MODE 0: FP32; 1: Kahan-compensated block sums; 2: FP16 far field.
template <int MODE>
struct WalkMode {
  static constexpr bool comp = MODE == 1;
  static constexpr bool half = MODE == 2;
  static constexpr int stack = kWaveStack;                 // LDS stack entries per wave
  static constexpr int l16 = half ? 2 * kHalfList : 1;     // half2 slots of the FP16 list
};

template <bool COUNT, int MODE, bool POT = false, int STACK = kWaveStack>
__device__ __forceinline__ void walk_group64(
    int t0, int count, int root, int slot, const float4* __restrict__ pos_sorted,
    const unsigned* __restrict__ perm, int n, const Node* __restrict__ nodes, float theta2,
    float eps2, float G, float4* __restrict__ acc, float4* __restrict__ vel, float half_dt,
    int* __restrict__ counts, int* __restrict__ opens, int* overflow, int* error, int* stk,
    float4* list, half2_t* l16 = nullptr, float* __restrict__ phi = nullptr) {
  constexpr bool COMP = WalkMode<MODE>::comp;
  constexpr bool HALF = WalkMode<MODE>::half;
  constexpr int WSTACK = STACK;
  static_assert(STACK >= kStackReserve + 128, "stack below the depth-first reserve");
  const int lane = threadIdx.x & 63;
  const bool live = lane < count;
  const int t = t0 + (live ? lane : count - 1);
  const int nm1 = n - 1;
  const float4 b = pos_sorted[t];
  const unsigned long long live_mask = __builtin_amdgcn_ballot_w64(live);

  // bounding box of the live targets (dead lanes repeat the last target)
  float blx = wave_min(b.x), bly = wave_min(b.y), blz = wave_min(b.z);
  float bhx = wave_max(b.x), bhy = wave_max(b.y), bhz = wave_max(b.z);
  const float acc_lo = theta2 * (1.f - kMacMargin);
  // FP16 far field: group-relative frame
  const float gcx = 0.5f * (blx + bhx), gcy = 0.5f * (bly + bhy), gcz = 0.5f * (blz + bhz);
  // per-axis room left for a cell's offset from the centre (see kHalfMaxSep)
  const float rmx = kHalfMaxSep - 0.5f * (bhx - blx), rmy = kHalfMaxSep - 0.5f * (bhy - bly),
              rmz = kHalfMaxSep - 0.5f * (bhz - blz);
  half2_t hbx = {}, hby = {}, hbz = {}, heps2 = {};
  float fx = 0.f, fy = 0.f, fz = 0.f;
  int n16 = 0;
  if (HALF) {
    hbx = (half2_t)((_Float16)(b.x - gcx));
    hby = (half2_t)((_Float16)(b.y - gcy));
    hbz = (half2_t)((_Float16)(b.z - gcz));
    heps2 = (half2_t)((_Float16)eps2);
  }
  // representative targets spread over the group (Morton order spreads them
  // in space); if one of them opens a node by its own canonical test, the
  // node is open -- an exact shortcut, no margin needed.  A dead lane holds
  // pos_sorted[n-1], a live member of this (last) group, so it is valid too.
  float rx[kGroupReps], ry[kGroupReps], rz[kGroupReps];
#pragma unroll
  for (int r = 0; r < kGroupReps; ++r) {
    const int src_lane = r * 63 / (kGroupReps - 1);
    rx[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b.x), src_lane));
    ry[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b.y), src_lane));
    rz[r] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b.z), src_lane));
  }

  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, pot = 0.f;
  int sp = 1, nlist = 0;
  long ni = 0, no = 0;
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  if (lane == 0) stk[0] = root;
  // exact-vote staging: the free upper half of the interaction list (fewer
  // than 64 entries are pending between flushes)
  float4* vbuf = list + 64;
  while (sp > 0) {
    // pop up to 64 entries; near the top of the stack fall back to one at a
    // time (a depth-first walk needs at most one entry per tree level, and
    // the reserve exceeds the deepest tree), so the stack cannot overflow
    int k = sp < 64 ? sp : 64;
    if (sp > WSTACK - kStackReserve) k = 1;
    // lane l < k takes stk[sp-1-l]; the other lanes re-read the last popped
    // entry (a valid id, so every load below stays in bounds) and are masked
    // off in every ballot
    const unsigned long long vmask = k >= 64 ? ~0ull : ((1ull << k) - 1ull);
    const int id = stk[max(sp - 1 - lane, sp - k)];
    sp -= k;
    pops += k;  // a walk over a valid tree pops each node at most once
    if ((__builtin_amdgcn_ballot_w64((unsigned)id >= nodes_total) & vmask) ||
        pops > nodes_total) {
      if (lane == 0) *error = 1;
      break;
    }
    const bool is_leaf = id >= nm1;
    const unsigned long long leaf_mask = __builtin_amdgcn_ballot_w64(is_leaf) & vmask;
    const unsigned long long node_mask = vmask & ~leaf_mask;
    // (x, y, z, m) of the leaf or node and the node's (left, right, size2),
    // loaded unconditionally (a leaf lane reads node 0's links, unused)
    const int nid = is_leaf ? 0 : id;
    const float4* srcp = is_leaf ? pos_sorted + (id - nm1) : &nodes[nid].com;
    const float4 src = *srcp;
    const int4 meta = *reinterpret_cast<const int4*>(&nodes[nid].left);
    const float s2 = __int_as_float(meta.z);
    // prefilter on every lane; only internal nodes' results are used
    const float ex = fmaxf(fmaxf(blx - src.x, src.x - bhx), 0.f);
    const float ey = fmaxf(fmaxf(bly - src.y, src.y - bhy), 0.f);
    const float ez = fmaxf(fmaxf(blz - src.z, src.z - bhz), 0.f);
    const float dmin2 = ex * ex + ey * ey + ez * ez;
    const bool sure_accept = s2 < acc_lo * dmin2;
    // all representative tests are evaluated (no short-circuit branches):
    // open if one of them has theta2*d2 <= s2, i.e. the smallest d2 decides
    float rmin = dist2(src.x - rx[0], src.y - ry[0], src.z - rz[0]);
#pragma unroll
    for (int r = 1; r < kGroupReps; ++r)
      rmin = fminf(rmin, dist2(src.x - rx[r], src.y - ry[r], src.z - rz[r]));
    const bool sure_open = !(s2 < theta2 * rmin);
    unsigned long long open_mask = __builtin_amdgcn_ballot_w64(sure_open) & node_mask;
    const unsigned long long amb =
        node_mask & ~open_mask & ~__builtin_amdgcn_ballot_w64(sure_accept);
    // exact vote for the nodes the prefilter could not settle: their (com,
    // size2) are staged in LDS and read back by LDS broadcast, four nodes per
    // iteration, then the remaining one to three one at a time
    if (amb) {
      const int j = mbcnt64(amb);  // an ambiguous lane's rank among the ambiguous
      const int na = __popcll(amb);
      if (__builtin_amdgcn_inverse_ballot_w64(amb)) vbuf[j] = make_float4(src.x, src.y, src.z, s2);
      unsigned long long opened = 0;  // bit q: the q-th ambiguous node opens
      int q = 0;
      for (; q + 4 <= na; q += 4) {
        const float4 v0 = vbuf[q], v1 = vbuf[q + 1], v2 = vbuf[q + 2], v3 = vbuf[q + 3];
        const bool o0 = !(v0.w < theta2 * dist2(v0.x - b.x, v0.y - b.y, v0.z - b.z));
        const bool o1 = !(v1.w < theta2 * dist2(v1.x - b.x, v1.y - b.y, v1.z - b.z));
        const bool o2 = !(v2.w < theta2 * dist2(v2.x - b.x, v2.y - b.y, v2.z - b.z));
        const bool o3 = !(v3.w < theta2 * dist2(v3.x - b.x, v3.y - b.y, v3.z - b.z));
        if (__builtin_amdgcn_ballot_w64(o0) & live_mask) opened |= 1ull << q;
        if (__builtin_amdgcn_ballot_w64(o1) & live_mask) opened |= 2ull << q;
        if (__builtin_amdgcn_ballot_w64(o2) & live_mask) opened |= 4ull << q;
        if (__builtin_amdgcn_ballot_w64(o3) & live_mask) opened |= 8ull << q;
      }
      for (; q < na; ++q) {
        const float4 v = vbuf[q];
        const bool o = !(v.w < theta2 * dist2(v.x - b.x, v.y - b.y, v.z - b.z));
        if (__builtin_amdgcn_ballot_w64(o) & live_mask) opened |= 1ull << q;
      }
      open_mask |= __builtin_amdgcn_ballot_w64((opened >> j) & 1ull) & amb;
    }
    const unsigned long long all_src = (node_mask & ~open_mask) | leaf_mask;
    // FP16 far field: accepted cells in range; everything else stays FP32
    unsigned long long far_mask = 0ull;
    if (HALF) {
      const bool safe16 = (src.w <= kHalfMaxAccel * (dmin2 + eps2)) && (src.w <= kHalfMaxMass) &&
                          (src.w >= kHalfMinMass) && (fabsf(src.x - gcx) < rmx) &&
                          (fabsf(src.y - gcy) < rmy) && (fabsf(src.z - gcz) < rmz);
      far_mask = node_mask & ~open_mask & __builtin_amdgcn_ballot_w64(safe16);
    }
    const unsigned long long src_mask = all_src & ~far_mask;
    if (COUNT) {
      ni += __popcll(all_src);
      no += __popcll(node_mask);
    }
    // push the children of opened nodes (lane order)
    const int nopen = __popcll(open_mask);
    if (sp + 2 * nopen > WSTACK) {
      if (lane == 0) *overflow = 1;
      break;
    }
    if (__builtin_amdgcn_inverse_ballot_w64(open_mask)) {
      const int ps = sp + 2 * mbcnt64(open_mask);
      stk[ps] = meta.x;
      stk[ps + 1] = meta.y;
    }
    sp += 2 * nopen;
    // append sources to the interaction list (lane order)
    if (__builtin_amdgcn_inverse_ballot_w64(src_mask)) list[nlist + mbcnt64(src_mask)] = src;
    nlist += __popcll(src_mask);
    if (HALF) {
      if (__builtin_amdgcn_inverse_ballot_w64(far_mask)) {
        const int j = n16 + mbcnt64(far_mask);
        _Float16* e = reinterpret_cast<_Float16*>(l16 + 4 * (j >> 1)) + (j & 1);
        e[0] = (_Float16)(src.x - gcx);
        e[2] = (_Float16)(src.y - gcy);
        e[4] = (_Float16)(src.z - gcz);
        e[6] = (_Float16)(src.w * kHalfMassScale);
      }
      n16 += __popcll(far_mask);
      if (n16 >= 64) {
        flush_half(l16, 32, hbx, hby, hbz, heps2, fx, fy, fz);
        n16 -= 64;
        // move the (< 64) remaining entries, pairs 32.., to the front
        l16[lane] = l16[128 + lane];
        l16[64 + lane] = l16[192 + lane];
      }
    }
    if (nlist >= 64) {
      flush_list<COMP, POT>(list, 64, b.x, b.y, b.z, eps2, ax, ay, az, cx, cy, cz, pot);
      nlist -= 64;
      if (lane < nlist) list[lane] = list[64 + lane];
    }
  }
  if (nlist > 0)
    flush_list<COMP, POT>(list, nlist, b.x, b.y, b.z, eps2, ax, ay, az, cx, cy, cz, pot);
  if (HALF) {
    if (n16 & 1) {  // pad the last pair with a massless entry
      if (lane < 4) reinterpret_cast<_Float16*>(l16 + 4 * (n16 >> 1))[2 * lane + 1] = 0;
    }
    if (n16 > 0) flush_half(l16, (n16 + 1) >> 1, hbx, hby, hbz, heps2, fx, fy, fz);
    constexpr float inv_scale = 1.f / kHalfMassScale;
    ax = fmaf(fx, inv_scale, ax);
    ay = fmaf(fy, inv_scale, ay);
    az = fmaf(fz, inv_scale, az);
  }
  if (COUNT && lane == 0) {
    counts[slot] = (int)ni;
    opens[slot] = (int)no;
  }
  if (POT && live) phi[t] = pot;
  if (live) finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

// Potential walk (tree energy): the wave64 walk at its own theta writing, per
// sorted target t, phi[t] = sum over its interaction list of m / sqrt(r^2 +
// eps^2) -- the target's own leaf included once (for theta^2 < 1/3 every node
// containing the target is opened).
__global__ __launch_bounds__(kBlock) void walk_potential_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float* __restrict__ phi,
    int* overflow, int* error) {
  constexpr int WAVES = kBlock / 64;
  __shared__ int s_stack[WAVES][kWaveStack];
  __shared__ float4 s_list[WAVES][kWaveList];
  const int w = threadIdx.x >> 6;
  const int gbase = xcd_block() * blockDim.x + (threadIdx.x & ~63);
  if (gbase >= n) return;  // wave-uniform
  walk_group64<false, 0, true>(gbase, min(64, n - gbase), 0, gbase >> 6, pos_sorted, perm, n,
                               nodes, theta2, eps2, 1.f, nullptr, nullptr, 0.f, nullptr, nullptr,
                               overflow, error, s_stack[w], s_list[w], nullptr, phi);
}

// Energy from the potential walk, FP64 per block: KE = sum m v^2 / 2,
// PE = -G/2 sum m (phi - m / eps), the self term removed with the same FP32
// value the walk added (m * rsq(eps^2)).
__global__ __launch_bounds__(kBlock) void tree_energy_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm,
    const float4* __restrict__ vel, const float* __restrict__ phi, int n, float eps2, float G,
    double* __restrict__ partial) {
  __shared__ double s_red[2][kBlock / 64];
  const int t = blockIdx.x * kBlock + threadIdx.x;
  double ke = 0.0, pe = 0.0;
  if (t < n) {
    const float4 b = pos_sorted[t];
    const unsigned o = perm[t];
    const float4 v = vel[o < (unsigned)n ? o : 0];
    const double m = (double)b.w;
    const double vx = v.x, vy = v.y, vz = v.z;
    ke = 0.5 * m * (vx * vx + vy * vy + vz * vz);
    const float self = b.w * rsq(eps2);
    pe = -0.5 * (double)G * m * ((double)phi[t] - (double)self);
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
    double a = 0.0, c = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) {
      a += s_red[0][k];
      c += s_red[1][k];
    }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = c;
  }
}

__global__ __launch_bounds__(kBlock) void tree_energy_final_kernel(
    const double* __restrict__ partial, int count, double* out) {
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
    double a = 0.0, c = 0.0;
    for (int k = 0; k < kBlock / 64; ++k) {
      a += s_red[0][k];
      c += s_red[1][k];
    }
    out[0] = a + c;
    out[1] = a;
    out[2] = c;
  }
}

// One group of `gs` (64, 32 or 16) consecutive targets per wave.  A group
// narrower than the wave runs the same node-parallel walk with the surplus
// lanes dead (they repeat the group's last target and vote with it): the
// visited / opened / accepted sets are exactly those of the gs-wide vote of
// traverse_warp_kernel, node batches still use all 64 lanes, and only the
// interaction-list evaluation idles the surplus lanes.
//
// STACK: LDS stack entries per wave.  The stack bounds occupancy (the kernel
// needs 72 VGPRs, room for 7 waves per SIMD; with 1024 entries a workgroup
// takes 24 KB of LDS and 6 fit a CU, with 768 entries 7); a shallower stack
// reaches the single-pop reserve sooner (1M two-disk, theta 0.5: the deepest
// group stack is 434 entries; theta 0.3: 756).
template <bool COUNT, int MODE, int STACK>
__global__ __launch_bounds__(kBlock) void walk_wave_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error, int gs) {
  constexpr int WAVES = kBlock / 64;
  __shared__ int s_stack[WAVES][STACK];
  __shared__ float4 s_list[WAVES][kWaveList];
  __shared__ half2_t s_l16[WAVES][WalkMode<MODE>::l16];
  const int w = threadIdx.x >> 6;
  const int group = xcd_block() * WAVES + w;
  const int gbase = group * gs;
  if (gbase >= n) return;  // wave-uniform
  walk_group64<COUNT, MODE, false, STACK>(gbase, min(gs, n - gbase), 0, group, pos_sorted, perm,
                                          n, nodes, theta2, eps2, G, acc, vel, half_dt, counts,
                                          opens, overflow, error, s_stack[w], s_list[w], s_l16[w]);
}

// Group widths 32 and 16: S = 64 / width subgroups of consecutive targets
// share one wave.  Each subgroup keeps the width-wide vote of
// traverse_warp_kernel (a node opens for it iff one of its live lanes opens
// the node) and its own interaction list, while the node batches are shared:
// a stack entry carries the set of subgroups that still visit the node, so a
// node is loaded and prefiltered once for all of them.  Every subgroup's
// visited / opened / accepted sets, hence its counts, are those of its own
// width-wide walk.  The S lists are evaluated together, each lane reading its
// subgroup's list, so the list work follows the longest list instead of the
// union of the lists that one 64-wide vote would evaluate.
constexpr float kPadCoord = 1e18f;  // massless list padding, far from every target
// The split walk runs half (width 32) or a quarter (width 16) as many waves as
// one group per wave, which loses while the grid does not fill the GPU: at
// width 32 one group per wave is faster up to 400K targets and slower from
// 600K on (1M: 1.29 -> 1.20 ms per step).
constexpr int kSplitMinN = 1 << 19;

__device__ __forceinline__ float read_lane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

template <bool COUNT, int S>
__device__ __forceinline__ void walk_split64(
    int t0, int count, const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm,
    int n, const Node* __restrict__ nodes, float theta2, float eps2, float G,
    float4* __restrict__ acc, float4* __restrict__ vel, float half_dt, int* __restrict__ counts,
    int* __restrict__ opens, int* overflow, int* error, int* stk, unsigned char* smask,
    float4* lists) {
  static_assert(S == 2 || S == 4, "subgroups of 32 or 16 targets");
  constexpr int W = 64 / S;
  constexpr int R = kReps / S;  // representative targets per subgroup
  constexpr unsigned long long kSubLanes = (1ull << W) - 1ull;
  const int lane = threadIdx.x & 63;
  const int sub = lane / W;
  const bool live = lane < count;
  const int t = t0 + (live ? lane : count - 1);
  const int nm1 = n - 1;
  const float4 b = pos_sorted[t];
  const unsigned long long live_mask = __builtin_amdgcn_ballot_w64(live);

  // subgroup boxes (xor offsets below W stay inside a subgroup; dead lanes
  // repeat the last live target, which lies in their own subgroup whenever
  // that subgroup is live), then every lane takes all S boxes and the
  // representatives of every subgroup
  float lx = b.x, ly = b.y, lz = b.z, hx = b.x, hy = b.y, hz = b.z;
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) {
    lx = fminf(lx, __shfl_xor(lx, o, 64));
    ly = fminf(ly, __shfl_xor(ly, o, 64));
    lz = fminf(lz, __shfl_xor(lz, o, 64));
    hx = fmaxf(hx, __shfl_xor(hx, o, 64));
    hy = fmaxf(hy, __shfl_xor(hy, o, 64));
    hz = fmaxf(hz, __shfl_xor(hz, o, 64));
  }
  float blx[S], bly[S], blz[S], bhx[S], bhy[S], bhz[S];
  float rx[S][R], ry[S][R], rz[S][R];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    blx[s] = read_lane(lx, s * W);
    bly[s] = read_lane(ly, s * W);
    blz[s] = read_lane(lz, s * W);
    bhx[s] = read_lane(hx, s * W);
    bhy[s] = read_lane(hy, s * W);
    bhz[s] = read_lane(hz, s * W);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int src_lane = s * W + r * (W - 1) / (R - 1);
      rx[s][r] = read_lane(b.x, src_lane);
      ry[s][r] = read_lane(b.y, src_lane);
      rz[s][r] = read_lane(b.z, src_lane);
    }
  }
  const float acc_lo = theta2 * (1.f - kMacMargin);

  float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, pot = 0.f;
  int sp = 1;
  int nl[S];
  long ni[S], no[S];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    nl[s] = 0;
    ni[s] = no[s] = 0;
  }
  const unsigned nodes_total = 2u * (unsigned)n - 1u;
  unsigned pops = 0;
  if (lane == 0) {
    stk[0] = 0;
    smask[0] = (unsigned char)((1u << ((count + W - 1) / W)) - 1u);  // the live subgroups
  }
  while (sp > 0) {
    // batches as in walk_group64; a node is pushed once, with the union of
    // the subgroups that opened its parent, so it is popped at most once
    int k = sp < 64 ? sp : 64;
    if (sp > kWaveStack - kStackReserve) k = 1;
    // lane l < k takes stk[sp-1-l]; the other lanes re-read the last popped
    // entry (in bounds) with no subgroup
    const bool valid = lane < k;
    const int slot = max(sp - 1 - lane, sp - k);
    const int id = stk[slot];
    const unsigned m = valid ? (unsigned)smask[slot] : 0u;
    sp -= k;
    pops += k;
    if (__ballot(valid && (unsigned)id >= nodes_total) || pops > nodes_total) {
      if (lane == 0) *error = 1;
      break;
    }
    const bool is_leaf = valid && id >= nm1;
    const bool is_node = valid && id < nm1;
    // (x, y, z, m) of the leaf or node and the node's (left, right, size2),
    // loaded unconditionally (other lanes read node 0, unused)
    const int nid = is_node ? id : 0;
    const float4* srcp = is_leaf ? pos_sorted + (id - nm1) : &nodes[nid].com;
    const float4 src = *srcp;
    const int4 meta = *reinterpret_cast<const int4*>(&nodes[nid].left);
    const float s2 = __int_as_float(meta.z);
    const int left = meta.x, right = meta.y;
    // per-subgroup prefilter on every lane, as in walk_group64 with the
    // subgroup's box and representatives; only internal nodes' bits are kept
    unsigned open_bits = 0, amb_bits = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const float ex = fmaxf(fmaxf(blx[s] - src.x, src.x - bhx[s]), 0.f);
      const float ey = fmaxf(fmaxf(bly[s] - src.y, src.y - bhy[s]), 0.f);
      const float ez = fmaxf(fmaxf(blz[s] - src.z, src.z - bhz[s]), 0.f);
      const float dmin2 = ex * ex + ey * ey + ez * ez;
      float rmin = dist2(src.x - rx[s][0], src.y - ry[s][0], src.z - rz[s][0]);
#pragma unroll
      for (int r = 1; r < R; ++r)
        rmin = fminf(rmin, dist2(src.x - rx[s][r], src.y - ry[s][r], src.z - rz[s][r]));
      const bool sure_accept = s2 < acc_lo * dmin2;
      const bool sure_open = !(s2 < theta2 * rmin);
      open_bits |= (unsigned)sure_open << s;
      amb_bits |= (unsigned)(!sure_accept & !sure_open) << s;
    }
    const unsigned node_bits = is_node ? m : 0u;
    open_bits &= node_bits;
    amb_bits &= node_bits;
    // exact vote for the unsettled (node, subgroup) pairs: every lane tests
    // each ambiguous node, four per iteration and then one at a time, and the
    // ballot is split by subgroup
    const unsigned long long amb = __builtin_amdgcn_ballot_w64(amb_bits != 0u);
    if (amb) {
      float4* vbuf = lists + 64;  // free upper half of list 0 (< 64 entries pending)
      const int j = mbcnt64(amb);
      if (amb_bits) vbuf[j] = make_float4(src.x, src.y, src.z, s2);
      const int na = __popcll(amb);
      unsigned long long opened[S];
#pragma unroll
      for (int s = 0; s < S; ++s) opened[s] = 0;
      int q = 0;
      for (; q + 4 <= na; q += 4) {
        unsigned long long bq[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 v = vbuf[q + i];
          bq[i] = __builtin_amdgcn_ballot_w64(
                      !(v.w < theta2 * dist2(v.x - b.x, v.y - b.y, v.z - b.z))) & live_mask;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int s = 0; s < S; ++s)
            if ((bq[i] >> (s * W)) & kSubLanes) opened[s] |= 1ull << (q + i);
      }
      for (; q < na; ++q) {
        const float4 v = vbuf[q];
        const unsigned long long bv = __builtin_amdgcn_ballot_w64(
                                          !(v.w < theta2 * dist2(v.x - b.x, v.y - b.y, v.z - b.z))) &
                                      live_mask;
#pragma unroll
        for (int s = 0; s < S; ++s)
          if ((bv >> (s * W)) & kSubLanes) opened[s] |= 1ull << q;
      }
#pragma unroll
      for (int s = 0; s < S; ++s)
        open_bits |= (((amb_bits >> s) & (unsigned)(opened[s] >> j)) & 1u) << s;
    }
    const unsigned acc_bits = is_leaf ? m : (node_bits & ~open_bits);
    if (COUNT) {
#pragma unroll
      for (int s = 0; s < S; ++s)
        no[s] += __popcll(__builtin_amdgcn_ballot_w64(is_node && ((m >> s) & 1u)));
    }
    // push the children of nodes some subgroup opened, with those subgroups
    const unsigned long long open_mask = __builtin_amdgcn_ballot_w64(open_bits != 0u);
    const int nopen = __popcll(open_mask);
    if (sp + 2 * nopen > kWaveStack) {
      if (lane == 0) *overflow = 1;
      break;
    }
    if (__builtin_amdgcn_inverse_ballot_w64(open_mask)) {
      const int ps = sp + 2 * mbcnt64(open_mask);
      stk[ps] = left;
      stk[ps + 1] = right;
      smask[ps] = (unsigned char)open_bits;
      smask[ps + 1] = (unsigned char)open_bits;
    }
    sp += 2 * nopen;
    // append each source to the list of every subgroup that accepts it
    int most = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const unsigned long long am = __builtin_amdgcn_ballot_w64((acc_bits >> s) & 1u);
      if (__builtin_amdgcn_inverse_ballot_w64(am)) lists[s * kWaveList + nl[s] + mbcnt64(am)] = src;
      nl[s] += __popcll(am);
      if (COUNT) ni[s] += __popcll(am);
      most = nl[s] > most ? nl[s] : most;
    }
    if (most >= 64) {
      // shorter lists are padded with massless entries (an exact +0 per
      // component), so all lanes evaluate 64 entries of their own list
      const float4 pad = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.f);
#pragma unroll
      for (int s = 0; s < S; ++s)
        if (lane >= nl[s]) lists[s * kWaveList + lane] = pad;
      flush_list<false, false>(lists + sub * kWaveList, 64, b.x, b.y, b.z, eps2, ax, ay, az, cx,
                               cy, cz, pot);
#pragma unroll
      for (int s = 0; s < S; ++s) {
        nl[s] = nl[s] > 64 ? nl[s] - 64 : 0;
        if (lane < nl[s]) lists[s * kWaveList + lane] = lists[s * kWaveList + 64 + lane];
      }
    }
  }
  int most = 0;
#pragma unroll
  for (int s = 0; s < S; ++s) most = nl[s] > most ? nl[s] : most;
  if (most > 0) {
    const float4 pad = make_float4(kPadCoord, kPadCoord, kPadCoord, 0.f);
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (lane >= nl[s] && lane < most) lists[s * kWaveList + lane] = pad;
    flush_list<false, false>(lists + sub * kWaveList, most, b.x, b.y, b.z, eps2, ax, ay, az, cx,
                             cy, cz, pot);
  }
  if (COUNT) {
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (lane == s * W && s * W < count) {
        counts[t0 / W + s] = (int)ni[s];
        opens[t0 / W + s] = (int)no[s];
      }
  }
  if (live) finish(t, n, ax, ay, az, G, perm, acc, vel, half_dt);
}

template <bool COUNT, int S>
__global__ __launch_bounds__(kBlock) void walk_split_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error) {
  constexpr int WAVES = kBlock / 64;
  __shared__ int s_stack[WAVES][kWaveStack];
  __shared__ unsigned char s_mask[WAVES][kWaveStack];
  __shared__ float4 s_list[WAVES][S * kWaveList];
  const int w = threadIdx.x >> 6;
  const int gbase = xcd_block() * blockDim.x + (threadIdx.x & ~63);
  if (gbase >= n) return;  // wave-uniform
  walk_split64<COUNT, S>(gbase, min(64, n - gbase), pos_sorted, perm, n, nodes, theta2, eps2, G,
                         acc, vel, half_dt, counts, opens, overflow, error, s_stack[w], s_mask[w],
                         s_list[w]);
}

// Persistent ensemble walk: each wave pulls work items (system, first sorted
// target, count <= 64) from a device queue until the queue is empty, walking
// the system's own subtree (root = roots[system]) of the forest.  Every wave
// exits once the counter passes the item count.
template <bool COUNT, int MODE>
__global__ __launch_bounds__(kBlock) void walk_items_kernel(
    const float4* __restrict__ pos_sorted, const unsigned* __restrict__ perm, int n,
    const Node* __restrict__ nodes, float theta2, float eps2, float G, float4* __restrict__ acc,
    float4* __restrict__ vel, float half_dt, int* __restrict__ counts, int* __restrict__ opens,
    int* overflow, int* error, const int4* __restrict__ items, int nitems,
    const int* __restrict__ roots, unsigned* queue) {
  constexpr int WAVES = kBlock / 64;
  __shared__ int s_stack[WAVES][WalkMode<MODE>::stack];
  __shared__ float4 s_list[WAVES][kWaveList];
  __shared__ half2_t s_l16[WAVES][WalkMode<MODE>::l16];
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  while (true) {
    unsigned k = 0;
    if (lane == 0) k = atomic_add_relaxed(queue, 1u);
    k = __builtin_amdgcn_readfirstlane(k);
    if (k >= (unsigned)nitems) break;
    const int4 it = items[k];  // (system, first target, count, -)
    walk_group64<COUNT, MODE>(it.y, it.z, roots[it.x], (int)k, pos_sorted, perm, n, nodes, theta2,
                              eps2, G, acc, vel, half_dt, counts, opens, overflow, error,
                              s_stack[w], s_list[w], s_l16[w]);
  }
}

#define NB_WALK_ARGS                                                                          \
  a.pos_sorted, a.perm, a.n, a.nodes, a.theta2, a.eps2, a.G, a.acc, a.vel, a.half_dt, a.counts, \
      a.opens, a.overflow, a.error

// The wave64 group walk in arithmetic MODE with the engine's stack depth.
template <bool COUNT, int MODE>
void launch_wave(hipStream_t s, int gw, int width, const WalkArgs& a) {
  switch (a.stack) {
    case 512:
      walk_wave_kernel<COUNT, MODE, 512><<<gw, kBlock, 0, s>>>(NB_WALK_ARGS, width);
      break;
    case 768:
      walk_wave_kernel<COUNT, MODE, 768><<<gw, kBlock, 0, s>>>(NB_WALK_ARGS, width);
      break;
    case 0:
    case 1024:
      walk_wave_kernel<COUNT, MODE, 1024><<<gw, kBlock, 0, s>>>(NB_WALK_ARGS, width);
      break;
    case 2048:
      walk_wave_kernel<COUNT, MODE, 2048><<<gw, kBlock, 0, s>>>(NB_WALK_ARGS, width);
      break;
    default:
      throw Error(kErrInvalid, "walk stack must be 512, 768, 1024 or 2048 entries");
  }
}

template <bool COUNT>
void walk_dispatch(hipStream_t s, int traverse, int width, const WalkArgs& a) {
  const int g = (a.n + kBlock - 1) / kBlock;
  if (traverse == 0) {
    walk_target_kernel<COUNT><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
  } else if (width == 64 || width == 32 || width == 16) {
    const int groups = (a.n + width - 1) / width;
    const int gw = (groups + kBlock / 64 - 1) / (kBlock / 64);
    if (a.half_far && !COUNT)
      launch_wave<COUNT, 2>(s, gw, width, a);
    else if (a.compensated && !COUNT)
      launch_wave<COUNT, 1>(s, gw, width, a);
    else if (width == 32 && a.n >= kSplitMinN)
      walk_split_kernel<COUNT, 2><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
    else if (width == 16 && a.n >= kSplitMinN)
      walk_split_kernel<COUNT, 4><<<g, kBlock, 0, s>>>(NB_WALK_ARGS);
    else
      launch_wave<COUNT, 0>(s, gw, width, a);
  } else {
    throw Error(kErrInvalid, "group width must be 16, 32 or 64");
  }
  NB_LAUNCH_CHECK();
}
#undef NB_WALK_ARGS

}  // namespace

void warm_traverse() {
  hipFuncAttributes a;
  NB_HIP(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(walk_wave_kernel<false, 0, 1024>)));
}

// Resident workgroups of the persistent walk: CUs x occupancy, queried once.
static int walk_items_grid() {
  static int grid = 0;
  if (grid == 0) {
    int dev = 0, cus = 0, per_cu = 0;
    NB_HIP(hipGetDevice(&dev));
    NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    NB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(walk_items_kernel<false, 0>), kBlock, 0));
    grid = std::max(1, cus) * std::max(1, per_cu);
  }
  return grid;
}

int walk_items_resident() { return walk_items_grid(); }

void launch_walk_items(hipStream_t s, const WalkArgs& a, const int4* items, int nitems,
                       const int* roots, unsigned* queue, int grid) {
  if (nitems <= 0) return;
  if (grid <= 0) grid = walk_items_grid();
  // a workgroup runs 4 waves; more workgroups than items/4 would only idle
  grid = std::min(grid, (nitems + kBlock / 64 - 1) / (kBlock / 64));
#define NB_ITEM_ARGS                                                                           \
  a.pos_sorted, a.perm, a.n, a.nodes, a.theta2, a.eps2, a.G, a.acc, a.vel, a.half_dt, a.counts, \
      a.opens, a.overflow, a.error, items, nitems, roots, queue
  if (a.counts)
    walk_items_kernel<true, 0><<<grid, kBlock, 0, s>>>(NB_ITEM_ARGS);
  else if (a.half_far)
    walk_items_kernel<false, 2><<<grid, kBlock, 0, s>>>(NB_ITEM_ARGS);
  else if (a.compensated)
    walk_items_kernel<false, 1><<<grid, kBlock, 0, s>>>(NB_ITEM_ARGS);
  else
    walk_items_kernel<false, 0><<<grid, kBlock, 0, s>>>(NB_ITEM_ARGS);
#undef NB_ITEM_ARGS
  NB_LAUNCH_CHECK();
}

int tree_energy_partials(int n) { return (n + kBlock - 1) / kBlock; }

void launch_tree_energy(hipStream_t s, const WalkArgs& a, float theta2, const float4* vel,
                        float* phi, double* partial, double* out) {
  if (a.n < 2) throw Error(kErrInvalid, "tree energy needs at least 2 particles");
  const int g = (a.n + kBlock - 1) / kBlock;
  walk_potential_kernel<<<g, kBlock, 0, s>>>(a.pos_sorted, a.perm, a.n, a.nodes, theta2, a.eps2,
                                             phi, a.overflow, a.error);
  NB_LAUNCH_CHECK();
  tree_energy_kernel<<<g, kBlock, 0, s>>>(a.pos_sorted, a.perm, vel, phi, a.n, a.eps2, a.G,
                                          partial);
  NB_LAUNCH_CHECK();
  tree_energy_final_kernel<<<1, kBlock, 0, s>>>(partial, g, out);
  NB_LAUNCH_CHECK();
}

void launch_walk(hipStream_t s, int traverse, int group_width, const WalkArgs& a) {
  if (a.n <= 0) return;
  if (a.counts)
    walk_dispatch<true>(s, traverse, group_width, a);
  else
    walk_dispatch<false>(s, traverse, group_width, a);
}

}  // namespace nb
