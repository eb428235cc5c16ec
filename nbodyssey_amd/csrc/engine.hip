// Host side of libnbody.so: the engine (device buffers, stream, timing), the
// C ABI of include/nbody.h and the reference-named C++ API of
// include/nbody/{simulation,forces,integrator}.hpp.
//
// Step structure (one leapfrog step, integrator.cu:27-37, all on one stream,
// no host synchronisation):
//   naive: kick_drift -> all-pairs (+ closing kick in its epilogue)
//   BH:    kick_drift_bbox (last block writes the FP64 quantisation and resets
//          the sort state) -> morton+histograms -> onesweep passes (the last
//          one gathers pos_sorted) -> Karras -> COM -> walk (+ closing kick)
#include "nb_internal.hpp"
#include "nbody.h"
#include "nbody/forces.hpp"
#include "nbody/integrator.hpp"
#include "nbody/simulation.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace nb {
namespace {

thread_local std::string g_last_error;

enum Phase { kPhBbox, kPhMorton, kPhSort, kPhBuild, kPhCom, kPhWalk, kPhases };
const char* kPhaseNames[kPhases] = {"bbox", "morton", "sort", "build", "com", "traverse"};

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t count = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void alloc(size_t c, bool zero = false) {
    reset();
    if (c == 0) c = 1;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&p), c * sizeof(T)));
    count = c;
    if (zero) {
      // the engine's own stream is non-blocking and does not wait for the
      // null stream, so finish the clear before anything else can touch p
      NB_HIP(hipMemset(p, 0, c * sizeof(T)));
      NB_HIP(hipDeviceSynchronize());
    }
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
};

void check_device() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    throw Error(kErrDevice, "no HIP device is visible (libnbody needs an MI355X / gfx950 GPU)");
}

// Loads every kernel code object once per process, before any timed call.
void warm_kernels() {
  static bool done = false;
  if (done) return;
  warm_integrate();
  warm_allpairs();
  warm_lbvh();
  warm_traverse();
  warm_ensemble();
  done = true;
}

}  // namespace

void set_last_error(const char* msg) { g_last_error = msg; }

void validate(const Params& p) {
  if (p.force != 0 && p.force != 1) throw Error(kErrInvalid, "force must be 0 (naive) or 1 (bh)");
  if (p.traverse != 0 && p.traverse != 1)
    throw Error(kErrInvalid, "traverse must be 0 (thread) or 1 (warp)");
  if (p.key_bits != 30 && p.key_bits != 63) throw Error(kErrInvalid, "key_bits must be 30 or 63");
  if (p.group_width != 16 && p.group_width != 32 && p.group_width != 64)
    throw Error(kErrInvalid, "group_width must be 16, 32 or 64");
  if (p.precision < 0 || p.precision > 2) throw Error(kErrInvalid, "precision must be 0, 1 or 2");
  if (p.precision == 2 && !(p.force == 1 && p.traverse == 1 && p.group_width == 64))
    throw Error(kErrInvalid,
                "precision 2 (FP16 far field) applies to the Barnes-Hut wave64 walk "
                "(force=bh, traverse=warp, group_width=64)");
  if (p.precision == 2 && !(p.eps >= 0.008f && p.eps <= 64.f))
    throw Error(kErrInvalid, "precision 2 (FP16 far field) needs 0.008 <= epsilon <= 64 "
                             "(eps^2 must be a normal FP16 number, and r^2 + eps^2 must stay "
                             "inside FP16 range)");
  if (!(p.theta >= 0.f)) throw Error(kErrInvalid, "theta must be >= 0");
  if (!(p.eps >= 0.f)) throw Error(kErrInvalid, "epsilon must be >= 0");
}

Params from_c(const nb_params* c) {
  Params p;
  p.dt = c->dt;
  p.eps = c->epsilon;
  p.G = c->G;
  p.theta = c->theta;
  p.force = c->force;
  p.traverse = c->traverse;
  p.key_bits = c->key_bits;
  p.group_width = c->group_width;
  p.precision = c->precision;
  return p;
}

BhScratch carve_bh_scratch(void* base, int n, size_t* bytes) {
  // 256-byte aligned pieces of one block, in the order of the fields
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t at = off;
    off += (b + 255) & ~size_t(255);
    return base ? static_cast<char*>(base) + at : nullptr;
  };
  const size_t c = (size_t)(n > 1 ? n : 1);
  BhScratch s{};
  // fixed-size bookkeeping first, then the per-particle arrays
  s.state = reinterpret_cast<SortState*>(take(sizeof(SortState)));
  s.quant = reinterpret_cast<QuantParams*>(take(sizeof(QuantParams)));
  s.bbox_partial = reinterpret_cast<float4*>(take(2 * (size_t)kBboxBlocks * sizeof(float4)));
  s.max_tiles = sort_max_tiles((int)c);
  s.lookback = reinterpret_cast<unsigned long long*>(
      take(lookback_words(s.max_tiles) * sizeof(unsigned long long)));
  s.keys_a = reinterpret_cast<uint64_t*>(take(c * sizeof(uint64_t)));
  s.keys_b = reinterpret_cast<uint64_t*>(take(c * sizeof(uint64_t)));
  s.vals_a = reinterpret_cast<unsigned*>(take(c * sizeof(unsigned)));
  s.vals_b = reinterpret_cast<unsigned*>(take(c * sizeof(unsigned)));
  s.pos_sorted = reinterpret_cast<float4*>(take(c * sizeof(float4)));
  s.nodes = reinterpret_cast<Node*>(take(c * sizeof(Node)));
  s.up = reinterpret_cast<int2*>(take(2 * c * sizeof(int2)));
  s.visit = reinterpret_cast<unsigned*>(take(c * sizeof(unsigned)));
  s.box_lo = reinterpret_cast<float4*>(take(c * sizeof(float4)));
  s.box_hi = reinterpret_cast<float4*>(take(c * sizeof(float4)));
  s.frontier = reinterpret_cast<int*>(take((size_t)com_scratch_ints((int)c) * sizeof(int)));
  if (bytes) *bytes = off;
  return s;
}

void enqueue_bh_force(hipStream_t stream, const BhScratch& b, float4* pos, float4* vel,
                      float4* acc, int n, const Params& p, bool step, PhaseMarks* marks,
                      int walk_stack) {
  const float half = 0.5f * p.dt;
  auto mark = [&](int k) {
    if (marks) marks->mark(k);
  };
  if (step)
    launch_kick_drift(stream, pos, vel, acc, n, half, p.dt, true, b.bbox_partial);
  else
    launch_bbox(stream, pos, n, b.bbox_partial);
  mark(0);
  launch_morton_hist(stream, pos, n, b.bbox_partial, bbox_parts(n), b.quant, p.key_bits,
                     sort_input(p.key_bits, b.keys_a, b.keys_b), b.state);
  mark(1);
  launch_sort(stream, p.key_bits, b.keys_a, b.keys_b, b.vals_a, b.vals_b, n, b.state, b.lookback,
              b.max_tiles, true, nullptr, nullptr);
  mark(2);
  launch_karras(stream, p.key_bits, b.keys_a, n, b.nodes, b.up, b.visit, b.state);
  mark(3);
  // the gather of the sorted positions runs in the first COM pass
  launch_com(stream, b.pos_sorted, n, b.nodes, b.up, b.visit, b.box_lo, b.box_hi, b.frontier,
             &b.state->error, b.vals_a, pos);
  mark(4);
  WalkArgs w{};
  w.pos_sorted = b.pos_sorted;
  w.perm = b.vals_a;
  w.n = n;
  w.nodes = b.nodes;
  w.theta2 = p.theta * p.theta;
  w.eps2 = p.eps * p.eps;
  w.G = p.G;
  w.acc = acc;
  w.vel = step ? vel : nullptr;
  w.half_dt = half;
  w.overflow = &b.state->overflow;
  w.error = &b.state->error;
  w.compensated = p.precision == 1;
  w.half_far = p.precision == 2;
  w.stack = walk_stack;
  launch_walk(stream, p.traverse, p.group_width, w);
  mark(5);
}

namespace {

// Events of one timed force call, resolved lazily (no host sync per call).
struct CallEvents {
  hipEvent_t ev[kPhases + 1];
  int count;  // 2 (total only) or kPhases + 1
  bool bh;
};

class Engine {
 public:
  explicit Engine(int n_) : n(n_) {
    check_device();
    warm_kernels();
    if (n < 1) throw Error(kErrInvalid, "particle count must be >= 1");
    NB_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_pos.alloc((size_t)n, true);
    own_vel.alloc((size_t)n, true);
    own_acc.alloc((size_t)n, true);
    pos = own_pos.p;
    vel = own_vel.p;
    acc = own_acc.p;
  }
  // Non-owning engine over caller device arrays, on the null stream (the
  // reference's default-stream semantics).
  Engine() : n(0) {
    check_device();
    warm_kernels();
  }
  ~Engine() {
    drop_graph();
    for (auto& c : pending_) release_events(c);
    for (auto e : free_events_) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  void bind(float4* p, float4* v, float4* a, int count) {
    if (count != n) acc_valid = false;
    if (p != pos || v != vel || a != acc || count != n) drop_graph();
    pos = p;
    vel = v;
    acc = a;
    n = count;
  }

  void set_params(const Params& np) {
    validate(np);
    if (nsys > 0 && (np.force != 1 || np.traverse != 1 || np.group_width != 64 ||
                     np.key_bits != 30))
      throw Error(kErrInvalid,
                  "an ensemble runs Barnes-Hut with the wave64 walk on (system, 30-bit) keys: "
                  "force=bh, traverse=warp, group_width=64, key_bits=30");
    if (np.key_bits != p.key_bits) release_bh();
    if (std::memcmp(&np, &p, sizeof(Params)) != 0) drop_graph();
    p = np;
  }

  // ---------------------------------------------------------------- forces
  // Enqueue one force evaluation on the current positions.  kick: this is
  // the force half of a leapfrog step, so the first kick + drift run before
  // it (fused with the bbox for BH) and the closing kick after it.
  void force_pass(bool step) {
    const float half = 0.5f * p.dt;
    const float eps2 = p.eps * p.eps;
    const bool bh = p.force == 1;
    CallEvents ce{};
    const bool timed = timing_on;
    // scratch allocation (first call, growth) happens before the timed region:
    // the force timing covers device work only
    prepare();
    if (timed) begin_call(ce, bh && phase_timing);
    if (nsys > 0) {
      ens_force_pass(step, ce);
    } else if (bh && n >= 2) {
      struct Marks : PhaseMarks {
        Engine* e;
        CallEvents* ce;
        void mark(int phase) override { e->mark(*ce, phase + 1); }
      } marks;
      marks.e = this;
      marks.ce = &ce;
      enqueue_bh_force(stream, scratch_view(), pos, vel, acc, n, p, step, &marks, walk_stack);
      tree_valid = true;
    } else if (bh) {
      // n < 2: zero force (forces_bh.cu:466-469)
      if (step) launch_kick_drift(stream, pos, vel, acc, n, half, p.dt, false, nullptr);
      launch_fill_u32(stream, reinterpret_cast<unsigned*>(acc), 0u, (size_t)n * 4);
      for (int k = 1; k <= kPhases; ++k) mark(ce, k);
    } else {
      if (step) launch_kick_drift(stream, pos, vel, acc, n, half, p.dt, false, nullptr);
      const int split = ap_split > 0 ? ap_split : allpairs_split(n);
      launch_allpairs(stream, pos, n, eps2, p.G, acc, ap_partial.p, split, p.precision == 1,
                      step ? vel : nullptr, half, ap_tile);
      mark(ce, 1);
    }
    if (timed) end_call(ce);
    acc_valid = true;
  }

  // One force evaluation over every system of the ensemble: the same pipeline
  // as a single system with (system, Morton) keys, a forest build and the
  // persistent walk over (system, target block) items.
  void ens_force_pass(bool step, CallEvents& ce) {
    const float half = 0.5f * p.dt;
    const int passes = ens_passes();
    launch_ens_kick_drift_bbox(stream, pos, vel, acc, n, half, p.dt, step, sys_off.p, nsys,
                               sys_box.p);
    mark(ce, kPhBbox + 1);
    launch_ens_morton_hist(stream, pos, n, sys_off.p, nsys, sys_box.p, keys_a.p, state.p);
    mark(ce, kPhMorton + 1);
    launch_sort_u64(stream, keys_a.p, keys_b.p, vals_a.p, vals_b.p, n, passes, state.p,
                    lookback.p, max_tiles, true, pos, pos_sorted.p);
    mark(ce, kPhSort + 1);
    launch_ens_karras(stream, keys_final(), n, sys_off.p, nsys, nodes.p, up.p, visit.p, roots.p,
                      state.p, sys_box.p);
    mark(ce, kPhBuild + 1);
    launch_com(stream, pos_sorted.p, n, nodes.p, up.p, visit.p, box_lo.p, box_hi.p, frontier.p,
               &state.p->error, nullptr, nullptr, sys_off.p, nsys);
    mark(ce, kPhCom + 1);
    WalkArgs w = walk_args();
    w.vel = step ? vel : nullptr;
    w.half_dt = half;
    launch_walk_items(stream, w, items.p, nitems, roots.p, &state.p->queue, walk_grid);
    mark(ce, kPhWalk + 1);
    tree_valid = true;
  }

  WalkArgs walk_args() {
    WalkArgs w{};
    w.pos_sorted = pos_sorted.p;
    w.perm = perm_final();
    w.n = n;
    w.nodes = nodes.p;
    w.theta2 = p.theta * p.theta;
    w.eps2 = p.eps * p.eps;
    w.G = p.G;
    w.acc = acc;
    w.overflow = &state.p->overflow;
    w.error = &state.p->error;
    w.compensated = p.precision == 1;
    w.half_far = p.precision == 2;
    w.stack = walk_stack;
    return w;
  }

  // Sorted keys / original indices after the last sort pass (the b buffers
  // for an odd number of passes).
  uint64_t* keys_final() { return (nsys > 0 && (ens_passes() & 1)) ? keys_b.p : keys_a.p; }
  unsigned* perm_final() { return (nsys > 0 && (ens_passes() & 1)) ? vals_b.p : vals_a.p; }

  // Turn this engine into an ensemble of ns systems of counts[s] particles
  // stored back to back.  Walk items: 64-target blocks, largest systems
  // first (the persistent walk then ends on the cheap items).
  void set_ensemble(const int* counts, int ns) {
    if (ns < 1 || ns > ens_max_systems())
      throw Error(kErrInvalid, "system count must be in [1, " +
                                   std::to_string(ens_max_systems()) + "]");
    std::vector<int> off(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
      if (counts[s] < 2) throw Error(kErrInvalid, "every system needs at least 2 particles");
      off[s + 1] = off[s] + counts[s];
    }
    if (off[ns] != n) throw Error(kErrInvalid, "system counts do not add up to the engine size");
    std::vector<int> order(ns);
    for (int s = 0; s < ns; ++s) order[s] = s;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return counts[a] > counts[b]; });
    std::vector<int4> it;
    for (int s : order)
      for (int t = off[s]; t < off[s + 1]; t += 64)
        it.push_back(make_int4(s, t, std::min(64, off[s + 1] - t), 0));
    std::vector<int2> tl;
    std::vector<int> toff(ns + 1, 0);
    for (int s = 0; s < ns; ++s) {
      for (int t = off[s]; t < off[s + 1]; t += kBlock) tl.push_back(make_int2(s, t));
      toff[s + 1] = (int)tl.size();
    }
    drop_graph();
    release_bh();
    sys_off.alloc(ns + 1);
    NB_HIP(hipMemcpy(sys_off.p, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice));
    items.alloc(it.size());
    NB_HIP(hipMemcpy(items.p, it.data(), it.size() * sizeof(int4), hipMemcpyHostToDevice));
    e_tiles.alloc(tl.size());
    NB_HIP(hipMemcpy(e_tiles.p, tl.data(), tl.size() * sizeof(int2), hipMemcpyHostToDevice));
    e_tile_off.alloc(ns + 1);
    NB_HIP(hipMemcpy(e_tile_off.p, toff.data(), toff.size() * sizeof(int), hipMemcpyHostToDevice));
    roots.alloc(ns, true);
    sys_box.alloc((size_t)8 * ns);
    nsys = ns;
    nitems = (int)it.size();
    ntiles_e = (int)tl.size();
    h_off = off;
    h_items = it;
    Params q = p;
    q.force = 1;
    q.traverse = 1;
    q.group_width = 64;
    q.key_bits = 30;
    p = q;
  }

  // (total, kinetic, potential) of every system, FP64.
  void system_energy(double* out) {
    if (nsys == 0) throw Error(kErrState, "not an ensemble");
    if ((int)en_partial.count < 2 * ntiles_e) en_partial.alloc((size_t)2 * ntiles_e);
    if ((int)en_out.count < 3 * nsys) en_out.alloc((size_t)3 * nsys);
    launch_ens_energy(stream, pos, vel, sys_off.p, e_tiles.p, ntiles_e, e_tile_off.p, nsys,
                      p.eps * p.eps, p.G, en_partial.p, en_out.p);
    NB_HIP(hipMemcpyAsync(out, en_out.p, 3 * (size_t)nsys * sizeof(double),
                          hipMemcpyDeviceToHost, stream));
    NB_HIP(hipStreamSynchronize(stream));
  }

  // Allocations a force pass needs, done before (never during) graph capture.
  void prepare() {
    if ((p.force == 1 && n >= 2) || nsys > 0) ensure_bh();
    if (p.force == 0) {
      const int split = ap_split > 0 ? ap_split : allpairs_split(n);
      if (split > 1 && ap_partial.count < (size_t)split * n) ap_partial.alloc((size_t)split * n);
    }
  }

  void step(int k) {
    if (k < 0) throw Error(kErrInvalid, "nsteps must be >= 0");
    if (k > 0 && !acc_valid) force_pass(false);
    if (!graph_mode) {
      for (int s = 0; s < k; ++s) force_pass(true);
      return;
    }
    if (k > 0 && !step_graph) capture_step();
    for (int s = 0; s < k; ++s) {
      CallEvents ce{};
      if (timing_on) begin_call(ce, false);
      NB_HIP(hipGraphLaunch(step_graph, stream));
      if (timing_on) end_call(ce);
    }
  }

  // One leapfrog step captured as a HIP graph: the same kernels with the
  // same arguments as force_pass(true); no memset nodes (the device-side
  // bookkeeping is reset by the kernels themselves), no events inside.
  void capture_step() {
    prepare();
    const bool saved = timing_on;
    timing_on = false;
    hipGraph_t g = nullptr;
    NB_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      force_pass(true);
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      if (g) (void)hipGraphDestroy(g);
      timing_on = saved;
      throw;
    }
    NB_HIP(hipStreamEndCapture(stream, &g));
    timing_on = saved;
    hipError_t e = hipGraphInstantiate(&step_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    NB_HIP(e);
  }

  void drop_graph() {
    if (step_graph) {
      (void)hipStreamSynchronize(stream);
      (void)hipGraphExecDestroy(step_graph);
      step_graph = nullptr;
    }
  }

  void check_overflow() {
    if (!state.p) return;
    int flag[2] = {0, 0};
    NB_HIP(hipMemcpyAsync(flag, &state.p->overflow, 2 * sizeof(int), hipMemcpyDeviceToHost,
                          stream));
    NB_HIP(hipStreamSynchronize(stream));
    if (flag[0] || flag[1]) {
      int zero[2] = {0, 0};
      NB_HIP(hipMemcpy(&state.p->overflow, zero, 2 * sizeof(int), hipMemcpyHostToDevice));
      if (flag[0])
        throw Error(kErrOverflow, "tree walk stack overflow (the walk needed more entries than "
                                  "its LDS stack holds)");
      throw Error(kErrState, "inconsistent sort/tree data detected on the device");
    }
  }

  void sync() {
    NB_HIP(hipStreamSynchronize(stream));
    resolve();
    check_overflow();
  }

  void energy(double out[3]) {
    if (n < 1) throw Error(kErrState, "no particles");
    if (nsys > 0) {  // no energy between systems
      std::vector<double> e((size_t)3 * nsys);
      system_energy(e.data());
      out[0] = out[1] = out[2] = 0.0;
      for (int s = 0; s < nsys; ++s)
        for (int k = 0; k < 3; ++k) out[k] += e[3 * s + k];
      return;
    }
    int np = energy_partials(n);
    if ((int)en_partial.count < 2 * np) en_partial.alloc((size_t)2 * np);
    if (en_out.count < 3) en_out.alloc(3);
    launch_energy(stream, pos, vel, n, p.eps * p.eps, p.G, en_partial.p, np, en_out.p);
    NB_HIP(hipMemcpyAsync(out, en_out.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    NB_HIP(hipStreamSynchronize(stream));
  }

  // ---------------------------------------------------------------- autotuner
  // Times candidate launch configurations on the current state and keeps the
  // fastest whose median force error against FP64 direct summation (256
  // fixed targets) stays within max_err; the particle state is saved and
  // restored, so tuning does not advance the simulation.
  //   Barnes-Hut: walk group width {64, 32, 16} x key bits {30, 63} x graph
  //               replay {off, on}, and at width 64 also the walk's LDS stack
  //               depth {512, 768, 1024, 2048} per wave;
  //   all-pairs:  source split {auto, 2x, 1/2, 1} x LDS tile width {256, 512,
  //               1024} x graph {off, on};
  //   ensemble:   persistent-walk grid {resident, 3/4, 1/2} x graph {off, on}.
  struct Candidate {
    int group_width, key_bits, graph, split, grid, tile, stack;
    double ms, err;
  };
  std::vector<Candidate> autotune(int steps, float max_err) {
    if (steps < 1) throw Error(kErrInvalid, "autotune needs steps >= 1");
    if (n < 2) throw Error(kErrInvalid, "autotune needs at least 2 particles");
    sync();
    const Params p0 = p;
    const bool graph0 = graph_mode;
    const int split0 = ap_split, grid0 = walk_grid, tile0 = ap_tile, stack0 = walk_stack;
    DevBuf<float4> save;
    save.alloc((size_t)3 * n);
    NB_HIP(hipMemcpyAsync(save.p, pos, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    NB_HIP(hipMemcpyAsync(save.p + n, vel, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
    NB_HIP(hipMemcpyAsync(save.p + 2 * (size_t)n, acc, n * sizeof(float4),
                          hipMemcpyDeviceToDevice, stream));
    const bool acc_valid0 = acc_valid;
    auto restore = [&]() {
      NB_HIP(hipMemcpyAsync(pos, save.p, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
      NB_HIP(hipMemcpyAsync(vel, save.p + n, n * sizeof(float4), hipMemcpyDeviceToDevice, stream));
      NB_HIP(hipMemcpyAsync(acc, save.p + 2 * (size_t)n, n * sizeof(float4),
                            hipMemcpyDeviceToDevice, stream));
      acc_valid = acc_valid0;
      tree_valid = false;
    };
    // fixed error targets, spread over the particles
    const int nt = std::min(n, 256);
    std::vector<int32_t> tg(nt);
    for (int k = 0; k < nt; ++k) tg[k] = (int32_t)(((long long)k * n) / nt);
    std::vector<double> ref((size_t)3 * nt);
    direct_f64(tg.data(), nt, ref.data());

    std::vector<Candidate> cands;
    if (nsys > 0) {
      const int full = walk_items_resident();
      for (int g : {full, full * 3 / 4, full / 2})
        for (int gr = 0; gr < 2; ++gr)
          cands.push_back({64, 30, gr, 0, std::max(g, 1), 0, 0, 0, 0});
    } else if (p.force == 1 && p.traverse == 0) {
      // the per-target walk keeps its own approximation: only key width and
      // graph replay are tuned (group widths would change the walk's result)
      for (int kb : {30, 63})
        for (int gr = 0; gr < 2; ++gr) cands.push_back({p.group_width, kb, gr, 0, 0, 0, 0, 0, 0});
    } else if (p.force == 1) {
      for (int kb : {30, 63})
        for (int w : {64, 32, 16})
          for (int gr = 0; gr < 2; ++gr) {
            cands.push_back({w, kb, gr, 0, 0, 0, 1024, 0, 0});
            if (w == 64)
              for (int st : {512, 768, 2048}) cands.push_back({w, kb, gr, 0, 0, 0, st, 0, 0});
          }
    } else {
      const int a = allpairs_split(n);
      std::vector<int> splits = {a, std::min(a * 2, 64), std::max(a / 2, 1), 1};
      std::sort(splits.begin(), splits.end());
      splits.erase(std::unique(splits.begin(), splits.end()), splits.end());
      for (int sp : splits)
        for (int tl : {256, 512, 1024})
          for (int gr = 0; gr < 2; ++gr)
            cands.push_back({p.group_width, p.key_bits, gr, sp, 0, tl, 0, 0, 0});
    }
    const bool timing0 = timing_on;
    hipEvent_t e0, e1;
    NB_HIP(hipEventCreate(&e0));
    NB_HIP(hipEventCreate(&e1));
    try {
      for (Candidate& c : cands) {
        Params q = p0;
        if (nsys == 0 && q.force == 1) {
          q.group_width = c.group_width;
          q.key_bits = c.key_bits;
        }
        set_params(q);
        ap_split = c.split;
        walk_grid = c.grid;
        ap_tile = c.tile;
        walk_stack = c.stack;
        drop_graph();
        graph_mode = c.graph != 0;
        timing_on = false;
        restore();
        // error of a(x) for this configuration
        force_pass(false);
        std::vector<float4> a(nt);
        std::vector<double> err(nt);
        for (int k = 0; k < nt; ++k)
          NB_HIP(hipMemcpyAsync(&a[k], acc + tg[k], sizeof(float4), hipMemcpyDeviceToHost, stream));
        NB_HIP(hipStreamSynchronize(stream));
        for (int k = 0; k < nt; ++k) {
          const double dx = a[k].x - ref[3 * k], dy = a[k].y - ref[3 * k + 1],
                       dz = a[k].z - ref[3 * k + 2];
          const double r = std::sqrt(ref[3 * k] * ref[3 * k] + ref[3 * k + 1] * ref[3 * k + 1] +
                                     ref[3 * k + 2] * ref[3 * k + 2]);
          err[k] = std::sqrt(dx * dx + dy * dy + dz * dz) / (r > 0 ? r : 1e-300);
        }
        std::nth_element(err.begin(), err.begin() + nt / 2, err.end());
        c.err = err[nt / 2];
        // time `steps` steps after one warmup step (captures the graph)
        step(1);
        NB_HIP(hipEventRecord(e0, stream));
        step(steps);
        NB_HIP(hipEventRecord(e1, stream));
        NB_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        NB_HIP(hipEventElapsedTime(&ms, e0, e1));
        c.ms = ms / steps;
        check_overflow();
      }
    } catch (...) {
      // leave the engine as the caller had it: parameters, launch choices and
      // the particle state (a candidate may have advanced it before failing)
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      timing_on = timing0;
      try {
        (void)hipStreamSynchronize(stream);
        set_params(p0);
        ap_split = split0;
        walk_grid = grid0;
        ap_tile = tile0;
        walk_stack = stack0;
        drop_graph();
        graph_mode = graph0;
        restore();
        (void)hipStreamSynchronize(stream);
      } catch (...) {
      }
      throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    timing_on = timing0;
    // fastest within the error bound; if none qualifies, the most accurate
    int best = -1;
    for (int k = 0; k < (int)cands.size(); ++k)
      if (cands[k].err <= max_err && (best < 0 || cands[k].ms < cands[best].ms)) best = k;
    if (best < 0) {
      best = 0;
      for (int k = 1; k < (int)cands.size(); ++k)
        if (cands[k].err < cands[best].err) best = k;
    }
    const Candidate& c = cands[best];
    Params q = p0;
    if (nsys == 0 && q.force == 1) {
      q.group_width = c.group_width;
      q.key_bits = c.key_bits;
    }
    set_params(q);
    ap_split = c.split;
    walk_grid = c.grid;
    ap_tile = c.tile;
    walk_stack = c.stack;
    drop_graph();
    graph_mode = c.graph != 0;
    restore();
    NB_HIP(hipStreamSynchronize(stream));
    std::swap(cands[0], cands[best]);  // the chosen one first
    return cands;
  }

  // Energy with the potential from a tree walk at theta_e (the O(n) cost
  // policy for large n): builds the tree for the current positions if the
  // last force pass did not, then walks it once more for the potential.
  void tree_energy(float theta_e, double out[3]) {
    if (nsys > 0) throw Error(kErrInvalid, "tree energy is for single systems (use the exact "
                                           "per-system energy of an ensemble)");
    if (p.force != 1) throw Error(kErrInvalid, "tree energy needs the Barnes-Hut force module");
    if (!(theta_e > 0.f && theta_e * theta_e < 1.f / 3.f))
      throw Error(kErrInvalid, "tree-energy theta must be in (0, 0.577)");
    if (!(p.eps > 0.f)) throw Error(kErrInvalid, "tree energy needs epsilon > 0");
    if (n < 2) throw Error(kErrInvalid, "tree energy needs at least 2 particles");
    if (!tree_valid) {
      // build the tree for the current positions; the accelerations of this
      // pass go to scratch, so the caller's acceleration array is untouched
      const bool saved = timing_on, saved_valid = acc_valid;
      float4* saved_acc = acc;
      if (scratch_acc.count < (size_t)n) scratch_acc.alloc((size_t)n);
      acc = scratch_acc.p;
      timing_on = false;
      try {
        force_pass(false);
      } catch (...) {
        acc = saved_acc;
        timing_on = saved;
        throw;
      }
      acc = saved_acc;
      acc_valid = saved_valid;
      timing_on = saved;
    }
    const int parts = tree_energy_partials(n);
    if (phi.count < (size_t)n) phi.alloc((size_t)n);
    if ((int)en_partial.count < 2 * parts) en_partial.alloc((size_t)2 * parts);
    if (en_out.count < 3) en_out.alloc(3);
    WalkArgs w = walk_args();
    launch_tree_energy(stream, w, theta_e * theta_e, vel, phi.p, en_partial.p, en_out.p);
    NB_HIP(hipMemcpyAsync(out, en_out.p, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    NB_HIP(hipStreamSynchronize(stream));
    check_overflow();
  }

  // FP64 direct-summation accelerations of the given targets (original
  // indices); an ensemble target's sources are its own system.
  void direct_f64(const int32_t* targets, int nt, double* acc3) {
    if (nt <= 0) return;
    std::vector<int> sys(nt, 0);
    for (int k = 0; k < nt; ++k) {
      if (targets[k] < 0 || targets[k] >= n) throw Error(kErrInvalid, "target index out of range");
      if (nsys > 0)
        sys[k] = (int)(std::upper_bound(h_off.begin(), h_off.end(), targets[k]) - h_off.begin()) - 1;
    }
    const int groups = nsys > 0 ? nsys : 1;
    std::vector<int> order(nt);
    for (int k = 0; k < nt; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sys[a] < sys[b]; });
    std::vector<int> tsorted(nt);
    for (int k = 0; k < nt; ++k) tsorted[k] = targets[order[k]];
    DevBuf<int> dt_;
    DevBuf<double> part, out;
    dt_.alloc(nt);
    NB_HIP(hipMemcpyAsync(dt_.p, tsorted.data(), nt * sizeof(int), hipMemcpyHostToDevice, stream));
    out.alloc((size_t)3 * nt);
    size_t need_part = 0;
    for (int g = 0, k = 0; g < groups && k < nt; ++g) {
      int k1 = k;
      while (k1 < nt && sys[order[k1]] == g) ++k1;
      if (k1 == k) continue;
      const int lo = nsys > 0 ? h_off[g] : 0, hi = nsys > 0 ? h_off[g + 1] : n;
      need_part = std::max(need_part, (size_t)direct_f64_split(k1 - k, hi - lo) * (k1 - k) * 3);
      k = k1;
    }
    part.alloc(need_part);
    const double eps2 = (double)p.eps * (double)p.eps;
    for (int g = 0, k = 0; g < groups && k < nt; ++g) {
      int k1 = k;
      while (k1 < nt && sys[order[k1]] == g) ++k1;
      if (k1 == k) continue;
      const int lo = nsys > 0 ? h_off[g] : 0, hi = nsys > 0 ? h_off[g + 1] : n;
      // one group at a time: the partial buffer is reused
      launch_direct_f64(stream, pos, lo, hi, dt_.p + k, k1 - k, eps2, (double)p.G, part.p,
                        out.p + (size_t)3 * k);
      NB_HIP(hipStreamSynchronize(stream));
      k = k1;
    }
    std::vector<double> res((size_t)3 * nt);
    NB_HIP(hipMemcpyAsync(res.data(), out.p, res.size() * sizeof(double), hipMemcpyDeviceToHost,
                          stream));
    NB_HIP(hipStreamSynchronize(stream));
    for (int k = 0; k < nt; ++k)
      for (int c = 0; c < 3; ++c) acc3[(size_t)3 * order[k] + c] = res[(size_t)3 * k + c];
  }

  // ---------------------------------------------------------------- scratch
  BhScratch scratch_view() {
    BhScratch b{};
    b.keys_a = keys_a.p;
    b.keys_b = keys_b.p;
    b.vals_a = vals_a.p;
    b.vals_b = vals_b.p;
    b.pos_sorted = pos_sorted.p;
    b.nodes = nodes.p;
    b.up = up.p;
    b.visit = visit.p;
    b.box_lo = box_lo.p;
    b.box_hi = box_hi.p;
    b.frontier = frontier.p;
    b.bbox_partial = bh_partial.p;
    b.quant = quant.p;
    b.state = state.p;
    b.lookback = lookback.p;
    b.max_tiles = max_tiles;
    return b;
  }
  void ensure_bh() {
    if (bh_cap >= n) return;
    release_bh();
    const size_t c = (size_t)n;
    keys_a.alloc(c);
    keys_b.alloc(c);
    vals_a.alloc(c);
    vals_b.alloc(c);
    pos_sorted.alloc(c);
    nodes.alloc(c);
    up.alloc(2 * c);
    visit.alloc(c);
    box_lo.alloc(c);
    box_hi.alloc(c);
    frontier.alloc((size_t)com_scratch_ints(n));
    // the COM list counters start at zero (each build's COM passes reset them)
    launch_fill_u32(stream, reinterpret_cast<unsigned*>(frontier.p), 0u, kComCounterInts);
    bh_partial.alloc(2 * kBboxBlocks);
    quant.alloc(1, true);
    state.alloc(1, true);
    max_tiles = sort_max_tiles(n);
    lookback.alloc(lookback_words(max_tiles), true);
    // the sort state is reset by each step's Karras kernel for the next step;
    // start from a reset state on this stream
    launch_sort_init(stream, state.p);
    if (nsys > 0) launch_ens_box_init(stream, sys_box.p, nsys);
    bh_cap = n;
  }
  void release_bh() {
    drop_graph();
    keys_a.reset(); keys_b.reset(); vals_a.reset(); vals_b.reset(); pos_sorted.reset();
    nodes.reset(); up.reset(); visit.reset(); box_lo.reset(); box_hi.reset(); frontier.reset();
    bh_partial.reset(); quant.reset(); state.reset(); lookback.reset();
    bh_cap = 0;
    tree_valid = false;
  }

  // ---------------------------------------------------------------- timing
  hipEvent_t get_event() {
    if (!free_events_.empty()) {
      hipEvent_t e = free_events_.back();
      free_events_.pop_back();
      return e;
    }
    hipEvent_t e;
    NB_HIP(hipEventCreate(&e));
    return e;
  }
  void release_events(CallEvents& c) {
    for (int k = 0; k < c.count; ++k) free_events_.push_back(c.ev[k]);
    c.count = 0;
  }
  void begin_call(CallEvents& c, bool phases) {
    c.bh = phases;
    c.count = phases ? kPhases + 1 : 2;
    for (int k = 0; k < c.count; ++k) c.ev[k] = get_event();
    NB_HIP(hipEventRecord(c.ev[0], stream));
  }
  void mark(CallEvents& c, int k) {
    if (!timing_on || !c.bh || k >= c.count - 1) return;
    NB_HIP(hipEventRecord(c.ev[k], stream));
  }
  void end_call(CallEvents& c) {
    NB_HIP(hipEventRecord(c.ev[c.count - 1], stream));
    pending_.push_back(c);
    if (pending_.size() > 512) resolve();
  }
  void resolve() {
    if (pending_.empty()) return;
    NB_HIP(hipEventSynchronize(pending_.back().ev[pending_.back().count - 1]));
    for (auto& c : pending_) {
      float ms = 0.f;
      NB_HIP(hipEventElapsedTime(&ms, c.ev[0], c.ev[c.count - 1]));
      force_ms += ms;
      ++force_calls;
      if (c.bh) {
        for (int k = 0; k < kPhases; ++k) {
          float pm = 0.f;
          NB_HIP(hipEventElapsedTime(&pm, c.ev[k], c.ev[k + 1]));
          phase_ms[k] += pm;
        }
        ++bh_calls;
      }
      release_events(c);
    }
    pending_.clear();
  }
  void reset_timing() {
    resolve();
    force_ms = 0.0;
    force_calls = 0;
    bh_calls = 0;
    for (double& x : phase_ms) x = 0.0;
  }

  int n;
  float4* pos = nullptr;
  float4* vel = nullptr;
  float4* acc = nullptr;
  hipStream_t stream = nullptr;
  Params p;
  bool acc_valid = false;
  bool tree_valid = false;
  bool timing_on = true;
  bool phase_timing = true;
  bool graph_mode = false;
  hipGraphExec_t step_graph = nullptr;
  double force_ms = 0.0;
  long force_calls = 0;
  double phase_ms[kPhases] = {};
  long bh_calls = 0;

  DevBuf<float4> own_pos, own_vel, own_acc;
  DevBuf<float> phi;
  DevBuf<float4> scratch_acc;
  DevBuf<float4> ap_partial;
  DevBuf<double> en_partial, en_out;
  int bh_cap = 0, max_tiles = 0;
  DevBuf<uint64_t> keys_a, keys_b;
  DevBuf<unsigned> vals_a, vals_b, visit;
  DevBuf<float4> pos_sorted, box_lo, box_hi, bh_partial;
  DevBuf<Node> nodes;
  DevBuf<int> frontier;
  DevBuf<int2> up;
  DevBuf<QuantParams> quant;
  DevBuf<SortState> state;
  DevBuf<unsigned long long> lookback;
  int ap_split = 0;    // all-pairs source split (0: automatic)
  int walk_grid = 0;   // persistent-walk workgroups (0: all resident)
  int ap_tile = 0;     // all-pairs sources per LDS tile (0: 256)
  int walk_stack = 0;  // wave64 group-walk LDS stack entries per wave (0: 1024)
  // ensemble (nsys > 0)
  int nsys = 0, nitems = 0, ntiles_e = 0;
  std::vector<int> h_off;
  std::vector<int4> h_items;
  DevBuf<int> sys_off, roots, e_tile_off;
  DevBuf<unsigned> sys_box;
  DevBuf<int4> items;
  DevBuf<int2> e_tiles;

 private:
  std::vector<hipEvent_t> free_events_;
  std::vector<CallEvents> pending_;
};

template <typename F>
int guarded(F&& f) {
  try {
    f();
    return kOk;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return kErrNoMem;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return kErrDevice;
  }
}

void need(const void* ptr, const char* what) {
  if (!ptr) throw Error(kErrInvalid, std::string(what) + " is NULL");
}

// Rejects a host float4 array holding NaN or infinity in one of its first
// `comps` components. Such a value would not fault a kernel (Morton keys are
// clamped and every walk is bounded), but it reaches every force through the
// tree's node masses and centres, so it is refused where the data enter.
void need_finite(const float* v, size_t count, int comps, const char* what) {
  const uint32_t kExp = 0x7f800000u;
  uint32_t bad = 0;
  for (size_t i = 0; i < count; ++i) {
    uint32_t u[4];
    std::memcpy(u, v + 4 * i, sizeof(u));
    for (int c = 0; c < comps; ++c) bad |= (u[c] & kExp) == kExp;
  }
  if (!bad) return;
  for (size_t i = 0; i < count; ++i)
    for (int c = 0; c < comps; ++c)
      if (!std::isfinite(v[4 * i + c]))
        throw Error(kErrInvalid, std::string(what) + " holds a non-finite value at particle " +
                                     std::to_string(i) + " component " + "xyzw"[c]);
}

// Host-buffer helpers for the stage-level entry points.
template <typename T>
void upload(DevBuf<T>& b, const T* h, size_t count) {
  b.alloc(count);
  NB_HIP(hipMemcpy(b.p, h, count * sizeof(T), hipMemcpyHostToDevice));
}

}  // namespace
}  // namespace nb

struct nb_engine {
  std::unique_ptr<nb::Engine> e;
};

using nb::guarded;
using nb::need;

extern "C" {

const char* nb_last_error(void) { return nb::g_last_error.c_str(); }
const char* nb_version(void) { return "nbody-mi355x 0.1 (gfx950)"; }

int nb_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

void nb_default_params(nb_params* p) {
  if (!p) return;
  p->dt = 0.01f;
  p->epsilon = 0.05f;
  p->G = 1.0f;
  p->theta = 0.5f;
  p->force = NB_FORCE_NAIVE;
  p->traverse = NB_TRAVERSE_THREAD;
  p->key_bits = 63;
  p->group_width = 32;
  p->precision = NB_PREC_FP32;
}

int nb_create(int n, nb_engine** out) {
  return guarded([&] {
    need(out, "out");
    *out = nullptr;
    auto h = std::make_unique<nb_engine>();
    h->e = std::make_unique<nb::Engine>(n);
    *out = h.release();
  });
}

int nb_create_ensemble(int nsys, const int* counts, nb_engine** out) {
  return guarded([&] {
    need(out, "out");
    need(counts, "counts");
    *out = nullptr;
    if (nsys < 1) throw nb::Error(nb::kErrInvalid, "nsys must be >= 1");
    long long total = 0;
    for (int s = 0; s < nsys; ++s) total += counts[s];
    if (total >= (1ll << 30)) throw nb::Error(nb::kErrInvalid, "ensemble too large (>= 2^30)");
    auto h = std::make_unique<nb_engine>();
    h->e = std::make_unique<nb::Engine>((int)std::max(total, 1ll));
    h->e->set_ensemble(counts, nsys);
    *out = h.release();
  });
}

int nb_ensemble_systems(const nb_engine* h) { return h ? h->e->nsys : -1; }

int nb_ensemble_offsets(const nb_engine* h, int32_t* offsets) {
  return guarded([&] {
    need(h, "engine");
    need(offsets, "offsets");
    if (h->e->nsys == 0) throw nb::Error(nb::kErrState, "not an ensemble");
    std::copy(h->e->h_off.begin(), h->e->h_off.end(), offsets);
  });
}

int nb_ensemble_items(const nb_engine* h, int32_t* items4, int* nitems) {
  return guarded([&] {
    need(h, "engine");
    if (h->e->nsys == 0) throw nb::Error(nb::kErrState, "not an ensemble");
    if (nitems) *nitems = h->e->nitems;
    if (items4)
      for (size_t k = 0; k < h->e->h_items.size(); ++k) {
        const int4 v = h->e->h_items[k];
        items4[4 * k] = v.x;
        items4[4 * k + 1] = v.y;
        items4[4 * k + 2] = v.z;
        items4[4 * k + 3] = v.w;
      }
  });
}

int nb_ensemble_roots(nb_engine* h, int32_t* roots) {
  return guarded([&] {
    need(h, "engine");
    need(roots, "roots");
    nb::Engine& e = *h->e;
    if (e.nsys == 0) throw nb::Error(nb::kErrState, "not an ensemble");
    if (!e.tree_valid) throw nb::Error(nb::kErrState, "no Barnes-Hut tree has been built");
    e.sync();
    NB_HIP(hipMemcpy(roots, e.roots.p, e.nsys * sizeof(int), hipMemcpyDeviceToHost));
  });
}

int nb_autotune(nb_engine* h, int steps, float max_median_err, nb_tune_result* out, char* log,
                int log_len) {
  return guarded([&] {
    need(h, "engine");
    std::vector<nb::Engine::Candidate> c = h->e->autotune(steps, max_median_err);
    if (out) {
      out->group_width = c[0].group_width;
      out->key_bits = c[0].key_bits;
      out->graph = c[0].graph;
      out->allpairs_split = c[0].split;
      out->walk_grid = c[0].grid;
      out->ms_per_step = c[0].ms;
      out->median_err = c[0].err;
      out->candidates = (int)c.size();
      out->allpairs_tile = c[0].tile;
      out->walk_stack = c[0].stack;
    }
    if (log && log_len > 0) {
      std::string t;
      char line[200];
      for (size_t k = 0; k < c.size(); ++k) {
        std::snprintf(line, sizeof line,
                      "%s width=%d keys=%d graph=%d split=%d tile=%d stack=%d grid=%d "
                      "ms/step=%.4f median_err=%.3e\n",
                      k == 0 ? "chosen " : "tried  ", c[k].group_width, c[k].key_bits, c[k].graph,
                      c[k].split, c[k].tile, c[k].stack, c[k].grid, c[k].ms, c[k].err);
        t += line;
      }
      std::snprintf(log, (size_t)log_len, "%s", t.c_str());
    }
  });
}

int nb_total_energy_tree(nb_engine* h, float theta, double out[3]) {
  return guarded([&] {
    need(h, "engine");
    need(out, "out");
    h->e->tree_energy(theta, out);
  });
}

int nb_direct_f64(nb_engine* h, const int32_t* targets, int nt, double* acc3) {
  return guarded([&] {
    need(h, "engine");
    if (nt < 0) throw nb::Error(nb::kErrInvalid, "nt must be >= 0");
    if (nt > 0) {
      need(targets, "targets");
      need(acc3, "acc3");
    }
    h->e->direct_f64(targets, nt, acc3);
  });
}

int nb_system_energy(nb_engine* h, double* out) {
  return guarded([&] {
    need(h, "engine");
    need(out, "out");
    h->e->system_energy(out);
  });
}

int nb_destroy(nb_engine* h) {
  return guarded([&] {
    if (!h) return;
    if (h->e && h->e->stream) (void)hipStreamSynchronize(h->e->stream);
    delete h;
  });
}

int nb_set_params(nb_engine* h, const nb_params* p) {
  return guarded([&] {
    need(h, "engine");
    need(p, "params");
    h->e->set_params(nb::from_c(p));
  });
}

int nb_get_params(const nb_engine* h, nb_params* out) {
  return guarded([&] {
    need(h, "engine");
    need(out, "params");
    const nb::Params& p = h->e->p;
    out->dt = p.dt;
    out->epsilon = p.eps;
    out->G = p.G;
    out->theta = p.theta;
    out->force = p.force;
    out->traverse = p.traverse;
    out->key_bits = p.key_bits;
    out->group_width = p.group_width;
    out->precision = p.precision;
  });
}

int nb_count(const nb_engine* h) { return h ? h->e->n : -1; }

int nb_upload(nb_engine* h, const float* pos4, const float* vel4) {
  return guarded([&] {
    need(h, "engine");
    need(pos4, "pos4");
    nb::Engine& e = *h->e;
    nb::need_finite(pos4, (size_t)e.n, 4, "pos4");
    if (vel4) nb::need_finite(vel4, (size_t)e.n, 3, "vel4");
    size_t bytes = (size_t)e.n * sizeof(float4);
    NB_HIP(hipMemcpyAsync(e.pos, pos4, bytes, hipMemcpyHostToDevice, e.stream));
    if (vel4)
      NB_HIP(hipMemcpyAsync(e.vel, vel4, bytes, hipMemcpyHostToDevice, e.stream));
    else
      NB_HIP(hipMemsetAsync(e.vel, 0, bytes, e.stream));
    NB_HIP(hipStreamSynchronize(e.stream));
    e.acc_valid = false;
    e.tree_valid = false;
  });
}

int nb_download(nb_engine* h, float* pos4, float* vel4) {
  return guarded([&] {
    need(h, "engine");
    nb::Engine& e = *h->e;
    size_t bytes = (size_t)e.n * sizeof(float4);
    if (pos4) NB_HIP(hipMemcpyAsync(pos4, e.pos, bytes, hipMemcpyDeviceToHost, e.stream));
    if (vel4) NB_HIP(hipMemcpyAsync(vel4, e.vel, bytes, hipMemcpyDeviceToHost, e.stream));
    e.sync();
  });
}

int nb_download_acc(nb_engine* h, float* acc4) {
  return guarded([&] {
    need(h, "engine");
    need(acc4, "acc4");
    nb::Engine& e = *h->e;
    NB_HIP(hipMemcpyAsync(acc4, e.acc, (size_t)e.n * sizeof(float4), hipMemcpyDeviceToHost,
                          e.stream));
    e.sync();
  });
}

int nb_compute_forces(nb_engine* h, float* acc4) {
  return guarded([&] {
    need(h, "engine");
    nb::Engine& e = *h->e;
    e.force_pass(false);
    if (acc4)
      NB_HIP(hipMemcpyAsync(acc4, e.acc, (size_t)e.n * sizeof(float4), hipMemcpyDeviceToHost,
                            e.stream));
    e.sync();
  });
}

int nb_step(nb_engine* h, int nsteps) {
  return guarded([&] {
    need(h, "engine");
    h->e->step(nsteps);
  });
}

int nb_synchronize(nb_engine* h) {
  return guarded([&] {
    need(h, "engine");
    h->e->sync();
  });
}

int nb_total_energy(nb_engine* h, double out[3]) {
  return guarded([&] {
    need(h, "engine");
    need(out, "out");
    h->e->energy(out);
  });
}

double nb_force_ms_total(nb_engine* h) {
  if (!h) return 0.0;
  if (guarded([&] { h->e->sync(); }) != 0) return -1.0;
  return h->e->force_ms;
}

long nb_force_calls(nb_engine* h) {
  if (!h) return 0;
  if (guarded([&] { h->e->sync(); }) != 0) return -1;
  return h->e->force_calls;
}

void nb_reset_force_timing(nb_engine* h) {
  if (!h) return;
  (void)guarded([&] { h->e->reset_timing(); });
}

int nb_set_phase_timing(nb_engine* h, int on) {
  return guarded([&] {
    need(h, "engine");
    h->e->resolve();
    h->e->phase_timing = on != 0;
  });
}

int nb_bh_phase_ms(nb_engine* h, double out[6], long* calls) {
  return guarded([&] {
    need(h, "engine");
    h->e->sync();
    if (out)
      for (int k = 0; k < nb::kPhases; ++k) out[k] = h->e->phase_ms[k];
    if (calls) *calls = h->e->bh_calls;
  });
}

int nb_set_graph(nb_engine* h, int on) {
  return guarded([&] {
    need(h, "engine");
    h->e->drop_graph();
    h->e->graph_mode = on != 0;
  });
}

int nb_set_launch(nb_engine* h, const nb_launch* l) {
  return guarded([&] {
    need(h, "engine");
    need(l, "launch");
    if (l->allpairs_split < 0 || l->allpairs_split > 64)
      throw nb::Error(nb::kErrInvalid, "allpairs_split must be in [0, 64]");
    if (l->allpairs_tile != 0 && l->allpairs_tile != 256 && l->allpairs_tile != 512 &&
        l->allpairs_tile != 1024)
      throw nb::Error(nb::kErrInvalid, "allpairs_tile must be 0, 256, 512 or 1024");
    if (l->walk_stack != 0 && l->walk_stack != 512 && l->walk_stack != 768 &&
        l->walk_stack != 1024 && l->walk_stack != 2048)
      throw nb::Error(nb::kErrInvalid, "walk_stack must be 0, 512, 768, 1024 or 2048");
    if (l->walk_grid < 0) throw nb::Error(nb::kErrInvalid, "walk_grid must be >= 0");
    nb::Engine& e = *h->e;
    e.drop_graph();
    e.ap_split = l->allpairs_split;
    e.ap_tile = l->allpairs_tile;
    e.walk_stack = l->walk_stack;
    e.walk_grid = l->walk_grid;
  });
}

int nb_get_launch(const nb_engine* h, nb_launch* l) {
  return guarded([&] {
    need(h, "engine");
    need(l, "launch");
    const nb::Engine& e = *h->e;
    l->allpairs_split = e.ap_split;
    l->allpairs_tile = e.ap_tile;
    l->walk_stack = e.walk_stack;
    l->walk_grid = e.walk_grid;
  });
}

int nb_bh_release(nb_engine* h) {
  return guarded([&] {
    need(h, "engine");
    NB_HIP(hipStreamSynchronize(h->e->stream));
    h->e->release_bh();
  });
}

int nb_time_steps(nb_engine* h, int nsteps, double* ms) {
  return guarded([&] {
    need(h, "engine");
    need(ms, "ms");
    nb::Engine& e = *h->e;
    if (nsteps > 0 && !e.acc_valid) e.force_pass(false);
    hipEvent_t a, b;
    NB_HIP(hipEventCreate(&a));
    NB_HIP(hipEventCreate(&b));
    // no per-call or per-phase events inside the bracket: each event record
    // leaves the GPU idle for a few microseconds between kernels
    const bool timing0 = e.timing_on;
    e.timing_on = false;
    NB_HIP(hipEventRecord(a, e.stream));
    try {
      e.step(nsteps);
    } catch (...) {
      e.timing_on = timing0;
      throw;
    }
    e.timing_on = timing0;
    NB_HIP(hipEventRecord(b, e.stream));
    NB_HIP(hipEventSynchronize(b));
    float t = 0.f;
    NB_HIP(hipEventElapsedTime(&t, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *ms = t;
    e.sync();
  });
}

int nb_get_tree(nb_engine* h, uint64_t* keys_sorted, uint32_t* perm, int32_t* children,
                int32_t* parent, float* com4, float* size2, float* pos_sorted4) {
  return guarded([&] {
    need(h, "engine");
    nb::Engine& e = *h->e;
    if (!e.tree_valid) throw nb::Error(nb::kErrState, "no Barnes-Hut tree has been built");
    e.sync();
    const int n = e.n;
    if (keys_sorted) {
      if (e.p.key_bits == 30 && e.nsys == 0) {
        std::vector<uint32_t> k(n);
        NB_HIP(hipMemcpy(k.data(), e.keys_a.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) keys_sorted[i] = k[i];
      } else {
        NB_HIP(hipMemcpy(keys_sorted, e.keys_final(), n * sizeof(uint64_t),
                         hipMemcpyDeviceToHost));
      }
    }
    if (perm)
      NB_HIP(hipMemcpy(perm, e.perm_final(), n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (parent) {
      std::vector<int2> u(2 * (size_t)n - 1);
      NB_HIP(hipMemcpy(u.data(), e.up.p, u.size() * sizeof(int2), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < u.size(); ++k) parent[k] = nb::up_parent(u[k]);
    }
    if (pos_sorted4)
      NB_HIP(hipMemcpy(pos_sorted4, e.pos_sorted.p, n * sizeof(float4), hipMemcpyDeviceToHost));
    if (children || com4 || size2) {
      std::vector<nb::Node> nd(n - 1);
      NB_HIP(hipMemcpy(nd.data(), e.nodes.p, (n - 1) * sizeof(nb::Node), hipMemcpyDeviceToHost));
      for (int i = 0; i < n - 1; ++i) {
        if (children) {
          children[2 * i] = nd[i].left;
          children[2 * i + 1] = nd[i].right;
        }
        if (com4) {
          com4[4 * i] = nd[i].com.x;
          com4[4 * i + 1] = nd[i].com.y;
          com4[4 * i + 2] = nd[i].com.z;
          com4[4 * i + 3] = nd[i].com.w;
        }
        if (size2) size2[i] = nd[i].size2;
      }
    }
  });
}

int nb_traverse_counts(nb_engine* h, int traverse, int group_width, int32_t* interactions,
                       int32_t* opens) {
  return guarded([&] {
    need(h, "engine");
    need(interactions, "interactions");
    nb::Engine& e = *h->e;
    if (!e.tree_valid) throw nb::Error(nb::kErrState, "no Barnes-Hut tree has been built");
    if (traverse != 0 && traverse != 1) throw nb::Error(nb::kErrInvalid, "bad traverse");
    if (traverse == 1 && group_width != 16 && group_width != 32 && group_width != 64)
      throw nb::Error(nb::kErrInvalid, "group_width must be 16, 32 or 64");
    const int n = e.n;
    if (e.nsys > 0 && (traverse != 1 || group_width != 64))
      throw nb::Error(nb::kErrInvalid, "an ensemble counts the wave64 walk only");
    const int slots = e.nsys > 0 ? e.nitems
                      : traverse == 0 ? n
                                      : (n + group_width - 1) / group_width;
    nb::DevBuf<int> ci, co;
    ci.alloc(slots, true);
    co.alloc(slots, true);
    nb::WalkArgs w = e.walk_args();
    w.acc = nullptr;
    w.compensated = false;
    w.half_far = false;
    w.counts = ci.p;
    w.opens = co.p;
    if (e.nsys > 0) {
      nb::launch_fill_u32(e.stream, &e.state.p->queue, 0u, 1);
      nb::launch_walk_items(e.stream, w, e.items.p, e.nitems, e.roots.p, &e.state.p->queue, 0);
    } else {
      nb::launch_walk(e.stream, traverse, group_width, w);
    }
    e.sync();
    NB_HIP(hipMemcpy(interactions, ci.p, slots * sizeof(int), hipMemcpyDeviceToHost));
    if (opens) NB_HIP(hipMemcpy(opens, co.p, slots * sizeof(int), hipMemcpyDeviceToHost));
  });
}

int nb_morton(const float* pos4, int n, int key_bits, uint64_t* keys) {
  return guarded([&] {
    need(pos4, "pos4");
    need(keys, "keys");
    if (n < 1) throw nb::Error(nb::kErrInvalid, "n must be >= 1");
    if (key_bits != 30 && key_bits != 63) throw nb::Error(nb::kErrInvalid, "key_bits must be 30 or 63");
    nb::check_device();
    nb::DevBuf<float4> dpos, part;
    nb::DevBuf<uint64_t> dkeys;
    nb::DevBuf<nb::QuantParams> qp;
    nb::DevBuf<nb::SortState> ss;
    nb::upload(dpos, reinterpret_cast<const float4*>(pos4), (size_t)n);
    dkeys.alloc(n);
    part.alloc(2 * nb::kBboxBlocks);
    qp.alloc(1, true);
    ss.alloc(1, true);
    nb::launch_bbox(nullptr, dpos.p, n, part.p);
    nb::launch_morton_hist(nullptr, dpos.p, n, part.p, nb::bbox_parts(n), qp.p, key_bits,
                           dkeys.p, ss.p);
    NB_HIP(hipDeviceSynchronize());
    if (key_bits == 30) {
      std::vector<uint32_t> k(n);
      NB_HIP(hipMemcpy(k.data(), dkeys.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
      for (int i = 0; i < n; ++i) keys[i] = k[i];
    } else {
      NB_HIP(hipMemcpy(keys, dkeys.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
    }
  });
}

int nb_sort_pairs(const uint64_t* keys, int n, int key_bits, uint64_t* keys_sorted,
                  uint32_t* perm) {
  return guarded([&] {
    need(keys, "keys");
    if (n < 1) throw nb::Error(nb::kErrInvalid, "n must be >= 1");
    if (key_bits != 30 && key_bits != 63) throw nb::Error(nb::kErrInvalid, "key_bits must be 30 or 63");
    nb::check_device();
    const int tiles = nb::sort_max_tiles(n);
    nb::DevBuf<uint64_t> ka, kb;
    nb::DevBuf<unsigned> va, vb;
    nb::DevBuf<nb::SortState> ss;
    nb::DevBuf<unsigned long long> lb;
    ka.alloc(n);
    kb.alloc(n);
    va.alloc(n);
    vb.alloc(n);
    ss.alloc(1, true);
    lb.alloc(nb::lookback_words(tiles), true);
    void* in = nb::sort_input(key_bits, ka.p, kb.p);
    if (key_bits == 30) {
      std::vector<uint32_t> k(n);
      for (int i = 0; i < n; ++i) {
        if (keys[i] >> 30) throw nb::Error(nb::kErrInvalid, "30-bit sort given a key >= 2^30");
        k[i] = (uint32_t)keys[i];
      }
      NB_HIP(hipMemcpy(in, k.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
      NB_HIP(hipMemcpy(in, keys, n * sizeof(uint64_t), hipMemcpyHostToDevice));
    }
    nb::launch_sort_init(nullptr, ss.p);
    nb::launch_hist(nullptr, in, n, key_bits, ss.p);
    nb::launch_sort(nullptr, key_bits, ka.p, kb.p, va.p, vb.p, n, ss.p, lb.p, tiles, true,
                    nullptr, nullptr);
    NB_HIP(hipDeviceSynchronize());
    if (keys_sorted) {
      if (key_bits == 30) {
        std::vector<uint32_t> k(n);
        NB_HIP(hipMemcpy(k.data(), ka.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) keys_sorted[i] = k[i];
      } else {
        NB_HIP(hipMemcpy(keys_sorted, ka.p, n * sizeof(uint64_t), hipMemcpyDeviceToHost));
      }
    }
    if (perm) NB_HIP(hipMemcpy(perm, va.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
  });
}

int nb_karras_build(const uint64_t* keys_sorted, int n, int key_bits, int32_t* children,
                    int32_t* parent) {
  return guarded([&] {
    need(keys_sorted, "keys_sorted");
    if (n < 2) throw nb::Error(nb::kErrInvalid, "n must be >= 2");
    if (key_bits != 30 && key_bits != 63) throw nb::Error(nb::kErrInvalid, "key_bits must be 30 or 63");
    nb::check_device();
    nb::DevBuf<uint64_t> dk;
    nb::DevBuf<nb::Node> nodes;
    nb::DevBuf<int2> par;
    nb::DevBuf<unsigned> visit;
    dk.alloc(n);
    if (key_bits == 30) {
      std::vector<uint32_t> k(n);
      for (int i = 0; i < n; ++i) k[i] = (uint32_t)keys_sorted[i];
      NB_HIP(hipMemcpy(dk.p, k.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    } else {
      NB_HIP(hipMemcpy(dk.p, keys_sorted, n * sizeof(uint64_t), hipMemcpyHostToDevice));
    }
    nodes.alloc(n - 1);
    par.alloc(2 * (size_t)n - 1);
    visit.alloc(n - 1);
    nb::launch_karras(nullptr, key_bits, dk.p, n, nodes.p, par.p, visit.p, nullptr);
    NB_HIP(hipDeviceSynchronize());
    if (parent) {
      std::vector<int2> u(2 * (size_t)n - 1);
      NB_HIP(hipMemcpy(u.data(), par.p, u.size() * sizeof(int2), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < u.size(); ++k) parent[k] = nb::up_parent(u[k]);
    }
    if (children) {
      std::vector<nb::Node> nd(n - 1);
      NB_HIP(hipMemcpy(nd.data(), nodes.p, (n - 1) * sizeof(nb::Node), hipMemcpyDeviceToHost));
      for (int i = 0; i < n - 1; ++i) {
        children[2 * i] = nd[i].left;
        children[2 * i + 1] = nd[i].right;
      }
    }
  });
}

}  // extern "C"

// ------------------------------------------------------------------------
// Reference-named C++ API (include/nbody/*.hpp).  One process-global engine
// bound to the caller's device arrays, on the null stream; errors abort with
// the HIP message, as the reference's CUDA_CHECK does.

namespace {

[[noreturn]] void die_api(const char* where, const std::exception& e) {
  std::fprintf(stderr, "HIP error in %s: %s\n", where, e.what());
  std::exit(EXIT_FAILURE);
}

template <typename F>
void api(const char* where, F&& body) {
  try {
    body();
  } catch (const std::exception& e) {
    die_api(where, e);
  }
}

nb::Engine* g_engine = nullptr;

nb::Engine& global_engine() {
  if (!g_engine) g_engine = new nb::Engine();
  return *g_engine;
}

nb::Params to_params(const SimParams& sp) {
  nb::Params p;
  p.dt = sp.dt;
  p.eps = sp.epsilon;
  p.G = sp.G;
  p.theta = sp.theta;
  p.force = sp.force == ForceMethod::kBarnesHut ? 1 : 0;
  p.traverse = sp.traverse == BhTraversal::kWarp ? 1 : 0;
  p.key_bits = sp.key_bits;
  p.group_width = sp.group_width;
  p.precision = sp.far16 ? 2 : (sp.compensated ? 1 : 0);
  return p;
}

void run_forces(const ParticleSystem& sys, float4* d_acc, const SimParams& sp, int force,
                bool timed) {
  nb::Engine& e = global_engine();
  nb::Params p = to_params(sp);
  p.force = force;
  if (force == 0 && p.precision == 2) p.precision = 0;  // the FP16 far field is a walk mode
  e.set_params(p);
  e.bind(sys.d_pos, sys.d_vel, d_acc, sys.n);
  e.timing_on = timed;
  e.force_pass(false);
  e.timing_on = true;
  e.sync();
}

}  // namespace

void ParticleSystem::allocate(int count) {
  api("ParticleSystem::allocate", [&] {
    if (count < 1) throw nb::Error(nb::kErrInvalid, "particle count must be >= 1");
    n = count;
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&d_pos), n * sizeof(float4)));
    NB_HIP(hipMalloc(reinterpret_cast<void**>(&d_vel), n * sizeof(float4)));
  });
}

void ParticleSystem::release() {
  api("ParticleSystem::release", [&] {
    if (d_pos) NB_HIP(hipFree(d_pos));
    if (d_vel) NB_HIP(hipFree(d_vel));
    d_pos = d_vel = nullptr;
    n = 0;
  });
}

void ParticleSystem::upload(const float4* h_pos, const float4* h_vel) {
  api("ParticleSystem::upload", [&] {
    NB_HIP(hipMemcpy(d_pos, h_pos, n * sizeof(float4), hipMemcpyHostToDevice));
    NB_HIP(hipMemcpy(d_vel, h_vel, n * sizeof(float4), hipMemcpyHostToDevice));
  });
}

void ParticleSystem::download_positions(float4* h_pos) const {
  api("ParticleSystem::download_positions", [&] { NB_HIP(hipMemcpy(h_pos, d_pos, n * sizeof(float4), hipMemcpyDeviceToHost)); });
}

void compute_forces(const ParticleSystem& sys, float4* d_acc, const SimParams& p) {
  api("compute_forces", [&] { run_forces(sys, d_acc, p, p.force == ForceMethod::kBarnesHut ? 1 : 0, true); });
}

void forces_naive(const ParticleSystem& sys, float4* d_acc, const SimParams& p) {
  api("forces_naive", [&] { run_forces(sys, d_acc, p, 0, false); });
}

void forces_barnes_hut(const ParticleSystem& sys, float4* d_acc, const SimParams& p) {
  api("forces_barnes_hut", [&] { run_forces(sys, d_acc, p, 1, false); });
}

double force_kernel_ms_total() {
  double v = 0.0;
  api("force_kernel_ms_total", [&] {
    global_engine().sync();
    v = global_engine().force_ms;
  });
  return v;
}

long force_kernel_calls() {
  long v = 0;
  api("force_kernel_calls", [&] {
    global_engine().sync();
    v = global_engine().force_calls;
  });
  return v;
}

void reset_force_timing() {
  api("reset_force_timing", [&] { global_engine().reset_timing(); });
}

void bh_report_phase_timing() {
  api("bh_report_phase_timing", [&] {
    nb::Engine& e = global_engine();
    e.sync();
    if (e.bh_calls == 0) return;
    double total = 0.0;
    for (int i = 0; i < nb::kPhases; ++i) total += e.phase_ms[i];
    std::printf("barnes-hut pipeline, per-call average over %ld calls:\n", e.bh_calls);
    for (int i = 0; i < nb::kPhases; ++i)
      std::printf("  %-9s %8.3f ms  (%4.1f%%)\n", nb::kPhaseNames[i], e.phase_ms[i] / e.bh_calls,
                  100.0 * e.phase_ms[i] / total);
  });
}

void bh_release() {
  api("bh_release", [&] {
    if (g_engine) {
      NB_HIP(hipDeviceSynchronize());
      g_engine->release_bh();
    }
  });
}

void leapfrog_step(ParticleSystem& sys, float4* d_acc, const SimParams& sp) {
  api("leapfrog_step", [&] {
    nb::Engine& e = global_engine();
    e.set_params(to_params(sp));
    e.bind(sys.d_pos, sys.d_vel, d_acc, sys.n);
    e.acc_valid = true;  // the caller's contract: d_acc holds a(x)
    e.force_pass(true);
  });
}

Energy total_energy_tree(const ParticleSystem& sys, const SimParams& sp, float theta_e) {
  Energy out{0.0, 0.0, 0.0};
  api("total_energy_tree", [&] {
    nb::Engine& e = global_engine();
    nb::Params p = to_params(sp);
    p.force = 1;
    if (p.traverse != 1 || p.group_width != 64 || p.precision == 2) {
      p.traverse = 1;
      p.group_width = 64;
      if (p.precision == 2) p.precision = 0;
    }
    e.set_params(p);
    e.bind(sys.d_pos, sys.d_vel, e.acc, sys.n);
    e.tree_valid = false;  // the caller may have changed the positions
    double r[3];
    e.tree_energy(theta_e, r);
    out = Energy{r[0], r[1], r[2]};
  });
  return out;
}

Energy total_energy(const ParticleSystem& sys, const SimParams& sp) {
  Energy out{0.0, 0.0, 0.0};
  api("total_energy", [&] {
    nb::Engine& e = global_engine();
    e.set_params(to_params(sp));
    e.bind(sys.d_pos, sys.d_vel, e.acc, sys.n);
    double r[3];
    e.energy(r);
    out = Energy{r[0], r[1], r[2]};
  });
  return out;
}
