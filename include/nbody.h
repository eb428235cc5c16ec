/*
 * nbody.h -- C ABI of libnbody.so, the MI355X (gfx950) N-body engine.
 *
 * Plain C types only (pointers, sizes, scalars).  Every function returns an
 * int status (NB_OK = 0) unless stated otherwise; nb_last_error() gives the
 * message of the last failure on the calling thread.  The reference aborts
 * the process instead (src/cuda_check.cuh:7-15); the galaxy_sim driver turns
 * a nonzero status back into the reference's message and exit(1).
 *
 * Host arrays use the reference's float4 layout (docs/FORMATS.md:11-16):
 * pos4 = n x (x, y, z, m), vel4 = n x (vx, vy, vz, 0), acc4 = n x (ax, ay, az, 0),
 * all in original particle order.  A handle owns its device state and one HIP
 * stream; one thread may use a handle at a time.
 *
 * The reference binds this path through a C++ API (src/forces.cuh,
 * src/integrator.cuh, src/simulation.hpp); the same-named C++ functions are
 * declared in include/nbody/simulation.hpp, forces.hpp and integrator.hpp.
 * Each entry point below names the reference interface it replaces.
 */
#ifndef NBODY_H
#define NBODY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_OK 0
#define NB_ERR_INVALID 1   /* bad argument */
#define NB_ERR_DEVICE 2    /* HIP runtime error / no device */
#define NB_ERR_NOMEM 3     /* device allocation failed */
#define NB_ERR_STATE 4     /* call made out of order */
#define NB_ERR_OVERFLOW 5  /* a tree-walk stack would have overflowed */

/* ForceMethod, src/simulation.hpp:4 */
#define NB_FORCE_NAIVE 0
#define NB_FORCE_BH 1
/* BhTraversal, src/simulation.hpp:5.  NB_TRAVERSE_WARP is the group-vote walk
 * of traverse_warp_kernel (src/forces_bh.cu:395-447) over `group_width` lanes
 * (16, 32 = the reference's warp, or 64 = one wave64). */
#define NB_TRAVERSE_THREAD 0
#define NB_TRAVERSE_WARP 1
/* arithmetic of the force sums:
 *   FP32          plain FP32 accumulation (the reference's);
 *   COMPENSATED   Kahan-added tile (all-pairs) / 64-source block (wave64 walk)
 *                 partial sums;
 *   FP16_FAR      wave64 walk only: accepted cells evaluated in packed FP16
 *                 (two interactions per instruction) relative to the group
 *                 centre with masses scaled by 512, products accumulated in
 *                 FP32; leaves and cells outside the safe FP16 range stay FP32.
 *                 Needs force BH, traverse WARP, group_width 64, eps >= 0.008. */
#define NB_PREC_FP32 0
#define NB_PREC_COMPENSATED 1
#define NB_PREC_FP16_FAR 2

/* SimParams, src/simulation.hpp:8-15, plus the additive knobs. */
typedef struct nb_params {
  float dt;        /* timestep */
  float epsilon;   /* softening length */
  float G;         /* gravitational constant */
  float theta;     /* Barnes-Hut opening angle; 0 = exact */
  int force;       /* NB_FORCE_* */
  int traverse;    /* NB_TRAVERSE_* */
  int key_bits;    /* Morton key width: 63 (reference) or 30 */
  int group_width; /* lanes per group for NB_TRAVERSE_WARP: 16, 32 or 64 */
  int precision;   /* NB_PREC_* */
} nb_params;

typedef struct nb_engine nb_engine;

/* Message of the last failed call on this thread ("" if none). */
const char* nb_last_error(void);
/* Library version string. */
const char* nb_version(void);
/* Number of visible HIP devices (0 when there is none); never fails. */
int nb_device_count(void);

/* Reference defaults: main.cu:19-25 and simulation.hpp:11-14 (dt 0.01,
 * eps 0.05, G 1, theta 0.5, naive, thread walk), key_bits 63, width 32. */
void nb_default_params(nb_params* p);

/* ParticleSystem::allocate (src/simulation.cu:4-8): device state for n
 * particles (n >= 1) on the current device. */
int nb_create(int n, nb_engine** out);
/* ParticleSystem::release + bh_release (src/simulation.cu:10-15,
 * src/forces_bh.cu:541-548). */
int nb_destroy(nb_engine* h);
int nb_set_params(nb_engine* h, const nb_params* p);
int nb_get_params(const nb_engine* h, nb_params* p);
int nb_count(const nb_engine* h);

/* ParticleSystem::upload (src/simulation.cu:17-20).  vel4 may be NULL (zero).
 * Returns NB_ERR_INVALID, copying nothing, when a position/mass component or
 * a velocity x/y/z is NaN or infinite (the reference copies such values). */
int nb_upload(nb_engine* h, const float* pos4, const float* vel4);
/* ParticleSystem::download_positions (src/simulation.cu:22-24); vel4 may be
 * NULL.  Waits for queued work. */
int nb_download(nb_engine* h, float* pos4, float* vel4);
/* Copies the engine's acceleration array (a(x) of the last force evaluation). */
int nb_download_acc(nb_engine* h, float* acc4);

/* compute_forces (src/forces.cuh:8, src/forces.cu:68-86): evaluates the force
 * module chosen by the params on the current positions into the engine's
 * acceleration array, then copies it to acc4 if acc4 is not NULL.  Timed into
 * nb_force_ms_total like the reference dispatcher. */
int nb_compute_forces(nb_engine* h, float* acc4);

/* leapfrog_step x nsteps (src/integrator.cuh:12-19, src/integrator.cu:27-37).
 * If no force evaluation has happened since the last upload, a(x0) is
 * computed first (main.cu:184).  Enqueued asynchronously; nb_synchronize or
 * any download waits. */
int nb_step(nb_engine* h, int nsteps);
int nb_synchronize(nb_engine* h);

/* total_energy (src/integrator.cuh:21, src/integrator.cu:93-110):
 * out = {total, kinetic, potential} reduced in FP64. */
int nb_total_energy(nb_engine* h, double out[3]);

/* force_kernel_ms_total / force_kernel_calls / reset_force_timing
 * (src/forces.cuh:17-19).  Wait for queued work. */
double nb_force_ms_total(nb_engine* h);
long nb_force_calls(nb_engine* h);
void nb_reset_force_timing(nb_engine* h);
/* Per-call phase timing off/on (default on).  Off removes the event records
 * between phases; the force-call total is still kept. */
int nb_set_phase_timing(nb_engine* h, int on);
/* bh_report_phase_timing data (src/forces_bh.cu:529-539): accumulated ms of
 * the six phases bbox, morton, sort, build, com, traverse and the call count. */
int nb_bh_phase_ms(nb_engine* h, double out[6], long* calls);
/* Graph mode off/on (default off).  On: nb_step replays one leapfrog step
 * captured as a HIP graph (captured on first use, re-captured after any
 * parameter change); results are bitwise identical to graph mode off.  Phase
 * timing is not recorded in graph mode (the force-call total is: one call per
 * step, kick and drift included).  No counterpart in the reference. */
int nb_set_graph(nb_engine* h, int on);
/* bh_release (src/forces_bh.cu:541-548): frees the tree scratch. */
int nb_bh_release(nb_engine* h);

/* Runs nsteps leapfrog steps bracketed by two HIP events on the engine stream
 * and returns the elapsed milliseconds (no host synchronisation inside).  No
 * per-call or per-phase events are recorded inside the bracket, so these
 * steps are not added to nb_force_ms_total / nb_force_calls / nb_bh_phase_ms. */
int nb_time_steps(nb_engine* h, int nsteps, double* ms);

/* ---- stage-level entry points (the reference keeps these stages in an
 * anonymous namespace, src/forces_bh.cu:39-461); used by the bit-exact
 * parity checks. */

/* Tree of the last Barnes-Hut evaluation: keys in sorted order (widened to
 * 64 bits), perm (original index of each sorted slot), children (2*(n-1)),
 * parent (2n-1), node COM (4*(n-1)), node size^2 (n-1), positions in sorted
 * order (4*n).  Any pointer may be NULL. */
int nb_get_tree(nb_engine* h, uint64_t* keys_sorted, uint32_t* perm, int32_t* children,
                int32_t* parent, float* com4, float* size2, float* pos_sorted4);
/* Runs the given walk on the current tree and returns its interaction and
 * opening-test counts: per sorted target for NB_TRAVERSE_THREAD, per group of
 * group_width consecutive sorted targets for NB_TRAVERSE_WARP.  Does not
 * touch the acceleration array. */
int nb_traverse_counts(nb_engine* h, int traverse, int group_width, int32_t* interactions,
                       int32_t* opens);
/* morton_kernel (src/forces_bh.cu:192-220) incl. the bbox reduction: keys of
 * n positions, key_bits 63 or 30. */
int nb_morton(const float* pos4, int n, int key_bits, uint64_t* keys);
/* cub::DeviceRadixSort::SortPairs (src/forces_bh.cu:490-497) with vals = index:
 * stable sort of keys (key_bits 30: keys < 2^30; 63: any 64-bit key). */
int nb_sort_pairs(const uint64_t* keys, int n, int key_bits, uint64_t* keys_sorted,
                  uint32_t* perm);
/* build_kernel (src/forces_bh.cu:242-277) on sorted keys. */
int nb_karras_build(const uint64_t* keys_sorted, int n, int key_bits, int32_t* children,
                    int32_t* parent);

/* total_energy (src/integrator.cu:43-91) at O(n log n) cost, the policy for
 * large n where the exact O(n^2) nb_total_energy takes minutes: KE exactly,
 * PE = -G/2 sum_i m_i phi_i with phi_i from a wave64 tree walk at `theta`
 * (in (0, 0.577), so every cell containing the target is opened and the self
 * term is removed exactly as the reference does, m_i / eps).  FP32 per-pair
 * potential, FP64 sums.  Needs the Barnes-Hut force module, epsilon > 0, a
 * single-system engine; builds the tree first if the last force evaluation
 * did not. */
int nb_total_energy_tree(nb_engine* h, float theta, double out[3]);

/* Launch configuration (no counterpart in the reference): kernel shapes
 * that change speed, not the physics.  0 selects the automatic value.
 *   allpairs_split  source partitions per 256-target block, [0, 64];
 *   allpairs_tile   sources per LDS tile of the all-pairs kernel: 256, 512
 *                   or 1024 (plain FP32 sums are the same bits at every
 *                   width; compensated sums Kahan-add per tile);
 *   walk_stack      LDS stack entries per wave of the wave64 group walk:
 *                   512, 768, 1024 (automatic) or 2048.  A shallower stack
 *                   lets more waves share a CU; a walk that nears its top
 *                   pops one node at a time (same result, slower);
 *   walk_grid       workgroups of the persistent ensemble walk (0: all that
 *                   fit the GPU at once).
 * Setting it drops a captured step graph (re-captured on the next step). */
typedef struct nb_launch {
  int allpairs_split;
  int allpairs_tile;
  int walk_stack;
  int walk_grid;
} nb_launch;
int nb_set_launch(nb_engine* h, const nb_launch* l);
int nb_get_launch(const nb_engine* h, nb_launch* l);

/* Autotuner (no counterpart in the reference): times candidate launch
 * configurations for `steps` leapfrog steps each on the current state, and
 * keeps the fastest whose median force error against FP64 direct summation
 * (256 fixed targets) is <= max_median_err (if none is, the most accurate).
 * Candidates: Barnes-Hut group walk -- group width 64/32/16 x key bits
 * 30/63 x graph replay off/on, and at width 64 the walk's LDS stack depth
 * 512/768/1024/2048; Barnes-Hut per-target walk (traverse THREAD, whose
 * approximation is kept) -- key bits x graph; all-pairs -- source split x
 * LDS tile width 256/512/1024 x graph; ensemble -- persistent-walk
 * workgroups x graph.  The particle state is saved and restored (tuning
 * does not advance the simulation), also when a candidate fails, in which
 * case the engine keeps the caller's parameters and launch configuration;
 * otherwise the choice stays set on the engine.  log (may be NULL) receives
 * one line per candidate, the chosen one first. */
typedef struct nb_tune_result {
  int group_width;
  int key_bits;
  int graph;           /* 1: steps replay a captured HIP graph */
  int allpairs_split;  /* 0: automatic */
  int walk_grid;       /* ensemble persistent-walk workgroups, 0: automatic */
  double ms_per_step;  /* of the chosen configuration */
  double median_err;   /* of the chosen configuration */
  int candidates;
  int allpairs_tile;   /* 0: automatic (256) */
  int walk_stack;      /* 0: automatic (1024) */
} nb_tune_result;
int nb_autotune(nb_engine* h, int steps, float max_median_err, nb_tune_result* out, char* log,
                int log_len);

/* FP64 direct-summation accelerations of nt targets (original particle
 * indices) from the engine's current positions: 3 doubles per target,
 * G * sum_{j != i} m_j (r_j - r_i) / (r^2 + eps^2)^{3/2} in FP64 on the
 * device.  The accuracy reference of the force modules: the FP64 form of
 * reference_nbody.py:45-61 (accelerations) that the reference's tests and
 * PERFORMANCE.md:84-92 measure Barnes-Hut error against.  In an ensemble the
 * sources of a target are its own system. */
int nb_direct_f64(nb_engine* h, const int32_t* targets, int nt, double* acc3);

/* ---------------------------------------------------------------- ensembles
 * No counterpart in the reference, which runs one system per process
 * (main.cu:72-240); additive, for batched runs of many independent systems.
 * The systems are stored back to back: system s holds particles
 * [offsets[s], offsets[s+1]) of the engine's arrays, so nb_upload,
 * nb_download, nb_compute_forces, nb_step and nb_set_graph work on all of
 * them at once.  Each step runs the single-system pipeline once over the
 * whole ensemble with (system << 30 | 30-bit Morton key) sort keys, a forest
 * build in which every system's subtree equals its standalone tree, and a
 * persistent wave64 walk pulling (system, 64-target block) items from a
 * device queue.  Parameters must be force = NB_FORCE_BH, traverse =
 * NB_TRAVERSE_WARP, group_width 64, key_bits 30; dt, epsilon, G, theta and
 * precision apply to every system. */

/* nsys in [1, 1024] systems of counts[s] >= 2 particles (total <= 2^30). */
int nb_create_ensemble(int nsys, const int* counts, nb_engine** out);
/* Number of systems (0 for a single-system engine, -1 for NULL). */
int nb_ensemble_systems(const nb_engine* h);
/* nsys + 1 offsets into the particle arrays. */
int nb_ensemble_offsets(const nb_engine* h, int32_t* offsets);
/* Walk work items in queue order, 4 ints each: (system, first sorted target,
 * target count <= 64, 0); *nitems receives their number.  items4 may be NULL.
 * nb_traverse_counts on an ensemble (traverse 1, width 64) reports one count
 * per item in this order. */
int nb_ensemble_items(const nb_engine* h, int32_t* items4, int* nitems);
/* Internal node id of each system's subtree root in the last forest built
 * (nb_get_tree returns the forest: sorted keys, perm, children, ...). */
int nb_ensemble_roots(nb_engine* h, int32_t* roots);
/* total_energy (src/integrator.cu:43-91) of every system on its own:
 * 3 doubles (total, kinetic, potential) per system.  nb_total_energy on an
 * ensemble returns the sums of these (no energy between systems). */
int nb_system_energy(nb_engine* h, double* out);

/* ------------------------------------------------------- no allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc -- the force evaluation of nb_compute_forces on
 * caller-owned device arrays, with no device allocation or free:
 *   d_pos4  n x (x, y, z, m) float4 in device memory of the current device;
 *   d_acc4  n x float4 device output (ax, ay, az, 0), original order;
 *   params  as for nb_set_params (force NAIVE or BH, any walk and precision);
 *   stream  a hipStream_t (NULL: the null stream); the work is enqueued on it
 *           and the call returns without waiting.
 * The result is bitwise that of nb_compute_forces with the same params and
 * positions.  Every scratch buffer lives in one static device array of
 * 12 MiB, sized for n <= NB_NOALLOC_MAX_N particles (larger n is rejected
 * with NB_ERR_INVALID), which each GPU the library runs this on holds until
 * the process ends.  All calls on one GPU share that array: calls that
 * overlap on one GPU (different streams or threads) race.  The walk's
 * consistency flags are not read back (that would need a host wait). */
#define NB_NOALLOC_MAX_N 65536
int nb_compute_forces_noalloc(const float* d_pos4, int n, const nb_params* params, float* d_acc4,
                              void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NBODY_H */
