// Internal declarations shared by the HIP translation units of libnbody.so.
//
// Device data layout (HBM):
//   pos, vel, acc        float4 x n, original particle order (simulation.hpp:17-28)
//   keys_a/keys_b        u32 (30-bit keys) or u64 (63-bit keys) x n, sort double buffer
//   vals_a/vals_b        u32 x n, original index riding with each key
//   pos_sorted           float4 x n, positions gathered into Morton order
//   nodes                Node (32 B) x (n-1), one record per internal node
//   up                   int2 x (2n-1): parent id | parent pass << 30, sibling (root: -1)
//   visit                u32 x (n-1), bottom-up arrival counters
//   box_lo/box_hi        float4 x (n-1), internal-node AABBs (used by the COM pass only)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace nb {

// Error raised inside the library; the C ABI turns it into a status code and
// a message (the reference aborts instead, cuda_check.cuh:7-15).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status {
  kOk = 0,
  kErrInvalid = 1,
  kErrDevice = 2,
  kErrNoMem = 3,
  kErrState = 4,
  kErrOverflow = 5,
};

#define NB_HIP(call)                                                                  \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      throw ::nb::Error(e_ == hipErrorOutOfMemory ? ::nb::kErrNoMem : ::nb::kErrDevice, \
                        std::string(#call) + ": " + hipGetErrorString(e_));           \
  } while (0)

#define NB_LAUNCH_CHECK() NB_HIP(hipGetLastError())

constexpr int kBlock = 256;          // kForceBlock (forces.cuh:25)
constexpr int kStackDepth = 128;     // kBhStack (forces_bh.cu:43)
constexpr int kBboxBlocks = 1024;    // max partial boxes written by the bbox / kick-drift-bbox pass
// Keys per thread of a onesweep tile, by key count: more, smaller tiles keep
// more workgroups per CU below 2^19 keys (100K 63-bit keys: 114 -> 94 us for
// 8 passes), 16 per thread up to 2^23 (1M: 103 us against 122 with 8), and
// 24 above (fewer tiles and look-back steps: 33.5M 63-bit keys 5.0 -> 4.2 ms).
constexpr int kSortItemsSmall = 8;
constexpr int kSortItems = 16;
constexpr int kSortItemsBig = 24;
inline int sort_items(int n) {
  return n >= (1 << 23) ? kSortItemsBig : n >= (1 << 19) ? kSortItems : kSortItemsSmall;
}
// Look-back rows needed by any sort of at most n keys (a buffer sized for n
// also serves smaller sorts, whose tiles may be smaller).
inline int sort_max_tiles(int n) {
  auto tiles = [](int m) { return (m + kBlock * sort_items(m) - 1) / (kBlock * sort_items(m)); };
  int t = tiles(n);
  if (n >= (1 << 19)) t = t > tiles((1 << 19) - 1) ? t : tiles((1 << 19) - 1);
  if (n >= (1 << 23)) t = t > tiles((1 << 23) - 1) ? t : tiles((1 << 23) - 1);
  return t;
}
constexpr int kMaxPasses = 8;
// Sort digit plan: 8-bit digits, 4 passes for 30-bit Morton keys, 8 for
// 63-bit keys and 5 for ensemble (system, 30-bit key) words below 2^40.
// Wider digits take fewer passes but each pass's look-back is slower (1M
// 30-bit keys: 3 x 10 bits 124 us, 4 x 8 bits 95.5 us; 4M 63-bit keys:
// 7 x 9 bits 563 us, 8 x 8 bits 541 us).  The onesweep kernel takes the
// digit width as a template parameter (up to 10 bits).
constexpr int kSortBits = 8;
// histogram words (pass p at p << bits), and look-back words per tile
constexpr int kHistWords = kMaxPasses << kSortBits;
// 64-bit look-back words of a sort buffer sized for max_tiles tiles: per pass,
// max_tiles rows of tile aggregates and max_tiles rows of group totals.
inline size_t lookback_words(int max_tiles) { return (size_t)2 * kHistWords * max_tiles; }
constexpr int sort_bits(int key_bits) { return key_bits > 0 ? kSortBits : 0; }
constexpr int sort_passes(int key_bits) { return key_bits == 30 ? 4 : 8; }
// The buffer a sort reads its unsorted keys from: with an odd pass count the
// passes start in the b buffers, so that the sorted result lands in a.
inline void* sort_input(int key_bits, void* keys_a, void* keys_b) {
  return (sort_passes(key_bits) & 1) ? keys_b : keys_a;
}

// One internal tree node, 32 bytes so a wave-uniform walk fetches it with a
// single scalar load.  com = (x, y, z, mass); size2 = (longest AABB edge)^2;
// internal node i covers the sorted leaves between i and `other` (inclusive).
struct alignas(32) Node {
  float4 com;
  int left;
  int right;
  float size2;
  int other;
};

// Morton quantisation constants, formed in FP64 by the bbox reduction.
struct QuantParams {
  double lo[3];
  double inv;
  float box_lo[4];
  float box_hi[4];
};

// Onesweep bookkeeping.  `epoch` tags look-back words so they never need
// clearing; `ticket` hands out tile ids in dispatch order; hist holds each
// pass's digit counts at pass * 2^bits.
struct SortState {
  unsigned hist[kHistWords];
  unsigned ticket[kMaxPasses];
  unsigned epoch;
  unsigned queue;           // work-item counter of the persistent ensemble walk
  int overflow;             // traversal stack overflow flag
  int error;                // consistency guard tripped (bad sort/tree data)
  int pad[3];
};

// Run parameters (SimParams of simulation.hpp:8-15 plus the additive knobs of
// nb_params); validate() throws kErrInvalid on an unsupported combination.
struct Params {
  float dt = 0.01f, eps = 0.05f, G = 1.0f, theta = 0.5f;
  int force = 0, traverse = 0, key_bits = 63, group_width = 32, precision = 0;
};
void validate(const Params& p);
// Message returned by nb_last_error() on this thread.
void set_last_error(const char* msg);

// ---------------------------------------------------------------- launchers
// Every launcher enqueues on `s` and checks the launch.

// k_integrate.hip
// The bbox passes write one (lo, hi) pair per workgroup into `partial`
// (bbox_parts(n) pairs); the Morton kernel reduces them.
int bbox_parts(int n);
void launch_kick_drift(hipStream_t s, float4* pos, float4* vel, const float4* acc, int n,
                       float half_dt, float dt, bool with_bbox, float4* partial);
void launch_bbox(hipStream_t s, const float4* pos, int n, float4* partial);
void launch_kick(hipStream_t s, float4* vel, const float4* acc, int n, float half_dt);
void launch_fill_u32(hipStream_t s, unsigned* p, unsigned v, size_t count);

// k_allpairs.hip
int allpairs_split(int n);
// split >= 1 source partitions; tile: sources per LDS tile, 256, 512 or 1024
// (0: 256).
void launch_allpairs(hipStream_t s, const float4* pos, int n, float eps2, float G,
                     float4* acc, float4* partial, int split, bool compensated,
                     float4* vel_kick, float half_dt, int tile = 0);
void launch_energy(hipStream_t s, const float4* pos, const float4* vel, int n, float eps2,
                   float G, double* partial, int max_partial, double* out);
int energy_partials(int n);
// FP64 direct sum for targets[0..nt) (original indices) over sources
// [lo, hi): out[3k..3k+2] = G * sum_{j != i} m_j d / (r2 sqrt(r2)).
// partial holds direct_f64_split(nt, hi - lo) * nt * 3 doubles.
int direct_f64_split(int nt, int nsrc);
void launch_direct_f64(hipStream_t s, const float4* pos, int lo, int hi, const int* targets,
                       int nt, double eps2, double G, double* partial, double* out);

// k_lbvh.hip
// Reduces the bbox partials, forms the FP64 quantisation (block 0 stores it
// in *qp), writes the keys and accumulates every pass's digit histogram.
// The sort state must be reset (launch_sort_init, or the previous step's
// Karras kernel via launch_karras(..., ss)).  The keys go to `keys`, which
// must be sort_input(key_bits, keys_a, keys_b) of the sort that follows.
void launch_morton_hist(hipStream_t s, const float4* pos, int n, const float4* partial,
                        int nparts, QuantParams* qp, int key_bits, void* keys, SortState* ss);
void launch_hist(hipStream_t s, const void* keys, int n, int key_bits, SortState* ss);
void launch_sort_init(hipStream_t s, SortState* ss);
// Sorts the keys in sort_input(key_bits, keys_a, keys_b) with their original
// indices (first_pass_identity: the values are the indices, not read from
// vals_*) through both buffer pairs; the result lands in keys_a / vals_a.  If
// gather_src is set, the last pass also writes gather_dst[dest] =
// gather_src[val].  The histograms must already be in ss (launch_morton_hist
// or launch_hist).  The sort sets SortState::error if a tile id or
// destination is out of range.
void launch_sort(hipStream_t s, int key_bits, void* keys_a, void* keys_b, unsigned* vals_a,
                 unsigned* vals_b, int n, SortState* ss, unsigned long long* lookback,
                 int max_tiles, bool first_pass_identity, const float4* gather_src,
                 float4* gather_dst);
// 64-bit keys sorted on their low passes * kSortBits bits; the result lands
// in the *_a buffers for an even pass count and in the *_b buffers for an
// odd one.  The histograms (pass p at p << kSortBits) must already be in ss.
void launch_sort_u64(hipStream_t s, uint64_t* keys_a, uint64_t* keys_b, unsigned* vals_a,
                     unsigned* vals_b, int n, int passes, SortState* ss,
                     unsigned long long* lookback, int max_tiles, bool first_pass_identity,
                     const float4* gather_src, float4* gather_dst);
// Also resets `ss` (histograms, tickets, epoch) for the next sort when ss is
// not null.
void launch_karras(hipStream_t s, int key_bits, const void* keys, int n, Node* nodes, int2* up,
                   unsigned* visit, SortState* ss);
// COM passes (k_lbvh.hip).  Pass 1 combines the nodes whose leaf range lies
// in one block of kComLeaves leaves, pass 2 those inside one superblock of
// kComSuper leaves, pass 3 the rest.  The Karras kernels store each internal
// node's pass in the top two bits of its `visit` word (com_visit_word); pass 2
// finds its nodes there and lists the pass-3 nodes it sees for pass 3.
constexpr int kComLeaves = 1024;
constexpr int kComSuper = 32768;
constexpr int kSeg = 256;         // frontier entries per workgroup segment
constexpr int kTopList = 4096;    // listed pass-3 nodes
// Scratch of the COM passes (ints, com_scratch_ints(n) of them; 16-byte
// aligned parts):
//   topcnt[kComCounterInts]    listed pass-3 nodes in word 0 (zero between builds)
//   toplist int2[kTopList]     (id, link) of the listed pass-3 nodes
//   seg1[g1 * kSeg], seg2[g2 * kSeg]   frontier entries (48 bytes: data, link, id)
//   rec1 int2[g1 * kSeg]       pass 1's records of later-pass nodes
//   cnt1[g1], rcnt1[g1], cnt2[g2]      segment counts
// with g1 = ceil(n / kComLeaves), g2 = ceil(n / kComSuper).
int com_scratch_ints(int n);
constexpr int kComCounterInts = 4;  // words at the start a new buffer must have zeroed
// The `visit` word of internal node i of leaf range [first, last]: its COM
// pass (0, 1 or 2) in bits 30-31, arrival bits clear.  Pass word 3 marks a
// node that gets no COM (an ensemble forest's nodes above the system roots).
__device__ __forceinline__ unsigned com_visit_word(int first, int last) {
  const unsigned pass = first / kComLeaves == last / kComLeaves ? 0u
                        : first / kComSuper == last / kComSuper ? 1u : 2u;
  return pass << 30;
}
// Upward record of a node or leaf (id: leaves n - 1 + index), written by its
// parent's Karras thread: x = parent id | the parent's `visit` pass word, or
// -1 at a root; y = sibling id.  The parent's pass is all the COM passes ask
// of its leaf range: pass 0 means it lies inside the child's pass-1 block,
// pass <= 1 inside the child's superblock.  Pass 2 sets the sign bit, so a
// root is x == -1 exactly (parent ids stay below 2^29, so no record is -1).
__device__ __forceinline__ int2 make_up(int parent, unsigned visit_word, int sibling) {
  return make_int2((int)((unsigned)parent | visit_word), sibling);
}
__device__ __forceinline__ int2 up_none() { return make_int2(-1, -1); }
__host__ __device__ inline bool up_root(int2 u) { return u.x == -1; }
__host__ __device__ inline int up_parent(int2 u) { return up_root(u) ? -1 : u.x & 0x3fffffff; }
__device__ __forceinline__ unsigned up_pass(int2 u) { return (unsigned)u.x >> 30; }
// With gather_src, the first pass also writes pos_sorted[i] =
// gather_src[perm[i]] (the sort then runs without its gather).  A forest
// (ensemble) passes its system offsets (nsys + 1 leaf offsets on the device):
// its systems' trees are independent, and the top pass splits over them.
void launch_com(hipStream_t s, float4* pos_sorted, int n, Node* nodes, const int2* up,
                unsigned* visit, float4* box_lo, float4* box_hi, int* scratch, int* error,
                const unsigned* perm = nullptr, const float4* gather_src = nullptr,
                const int* sys_off = nullptr, int nsys = 0);

// k_traverse.hip
// Each kernel translation unit's code object is loaded by the runtime at its
// first use; these resolve one kernel of each so that the loading happens when
// an engine is created rather than inside the first timed force call.
void warm_integrate();
void warm_allpairs();
void warm_lbvh();
void warm_traverse();
void warm_ensemble();

struct WalkArgs {
  const float4* pos_sorted;
  const unsigned* perm;
  int n;
  const Node* nodes;
  float theta2, eps2, G;
  float4* acc;        // output, original order (may be null in counting mode)
  float4* vel;        // if set: vel[perm[t]] += half_dt * acc (closing kick)
  float half_dt;
  int* counts;        // per target (thread walk) or per group: interactions
  int* opens;         // same, opening tests
  int* overflow;      // set to 1 if a stack would overflow
  int* error;         // set to 1 if a node id is out of range or the walk is too long
  bool compensated;   // Kahan-add each block of interactions (wave64 walk)
  bool half_far;      // FP16 far field for accepted cells (wave64 walk)
  int stack;          // LDS stack entries per wave of the wave64 group walk:
                      // 512, 768, 1024 or 2048 (0: 1024)
};
void launch_walk(hipStream_t s, int traverse, int group_width, const WalkArgs& a);
// Energy from a potential walk at theta2 (< 1/3) on the built tree:
// phi (float x n, sorted order) and partial (2 x tree_energy_partials(n)
// doubles) are scratch; out = (total, kinetic, potential), FP64.
int tree_energy_partials(int n);
void launch_tree_energy(hipStream_t s, const WalkArgs& a, float theta2, const float4* vel,
                        float* phi, double* partial, double* out);
// Persistent wave64 walk over ensemble work items (system, first target,
// count <= 64, 0), each system from its own subtree root; `queue` must be 0.
void launch_walk_items(hipStream_t s, const WalkArgs& a, const int4* items, int nitems,
                       const int* roots, unsigned* queue, int grid);
// Workgroups of the persistent walk that are resident at once (CUs x occupancy);
// grid <= 0 above means this many.
int walk_items_resident();

// k_ensemble.hip: systems stored back to back, system s at [off[s], off[s+1]).
// box: 8 u32 per system (orderable-float min xyz, max xyz, 2 unused), reset
// to empty by launch_ens_box_init and afterwards by each step's Karras kernel.
int ens_passes();
int ens_max_systems();
void launch_ens_box_init(hipStream_t s, unsigned* box, int nsys);
void launch_ens_kick_drift_bbox(hipStream_t s, float4* pos, float4* vel, const float4* acc,
                                int n, float half_dt, float dt, bool kick, const int* off,
                                int nsys, unsigned* box);
void launch_ens_morton_hist(hipStream_t s, const float4* pos, int n, const int* off, int nsys,
                            const unsigned* box, uint64_t* keys, SortState* ss);
// Also resets the sort state, the walk queue and the
// boxes for the next step, and writes roots[s] = the internal node whose leaf
// range is exactly system s.
void launch_ens_karras(hipStream_t s, const uint64_t* keys, int n, const int* off, int nsys,
                       Node* nodes, int2* up, unsigned* visit, int* roots, SortState* ss,
                       unsigned* box);
// tiles: (system, first target) per 256-target tile; tile_off[s] = first tile
// of system s; out: (total, kinetic, potential) per system.
void launch_ens_energy(hipStream_t s, const float4* pos, const float4* vel, const int* off,
                       const int2* tiles, int ntiles, const int* tile_off, int nsys, float eps2,
                       float G, double* partial, double* out);

// ---------------------------------------------------------------- BH pass
// Device buffers of one Barnes-Hut force evaluation over n particles.
struct BhScratch {
  uint64_t* keys_a;     // n keys (u32 words for 30-bit keys), sort double buffer
  uint64_t* keys_b;
  unsigned* vals_a;     // n original indices; after the sort: perm
  unsigned* vals_b;
  float4* pos_sorted;   // n
  Node* nodes;          // n - 1
  int2* up;             // 2n - 1
  unsigned* visit;      // n - 1
  float4* box_lo;       // n - 1
  float4* box_hi;       // n - 1
  int* frontier;        // com_scratch_ints(n)
  float4* bbox_partial; // 2 * kBboxBlocks
  QuantParams* quant;
  SortState* state;     // reset by launch_sort_init before the first pass
  unsigned long long* lookback;  // lookback_words(max_tiles)
  int max_tiles;
};
// The buffers for n particles carved out of one block at `base` in 256-byte
// aligned pieces, the fixed-size ones (state, quant, bbox_partial, lookback)
// first and keys_a the first per-particle piece; *bytes receives the block
// size (base may be null to size it).  A carve for n also serves any smaller
// particle count.
BhScratch carve_bh_scratch(void* base, int n, size_t* bytes);

// Phase hook: mark(k) is called after phase k (bbox, morton, sort, build, com,
// traverse) has been enqueued.
struct PhaseMarks {
  virtual void mark(int phase) = 0;
};
// One Barnes-Hut force evaluation (n >= 2) on `pos` into `acc` (original
// order), enqueued on `stream` (forces_bh.cu:463-527).  step: the first kick
// and the drift run first (fused with the bbox) and the closing kick rides in
// the walk's epilogue on `vel`.  The sort state must have been reset (by
// launch_sort_init, or by the previous pass's Karras kernel).  walk_stack:
// LDS stack entries per wave of the wave64 group walk (WalkArgs::stack).
void enqueue_bh_force(hipStream_t stream, const BhScratch& b, float4* pos, float4* vel,
                      float4* acc, int n, const Params& p, bool step, PhaseMarks* marks,
                      int walk_stack = 0);

}  // namespace nb

// C ABI parameter block (include/nbody.h).
struct nb_params;
namespace nb {
Params from_c(const ::nb_params* c);
}  // namespace nb
