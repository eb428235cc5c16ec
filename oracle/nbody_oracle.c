/*
 * nbody_oracle.c -- CPU restatement of the reference N-body hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product path (nbodyssey_amd/libnbody.so)
 * never links, loads or calls it.
 *
 * Every function restates one piece of the reference's algorithm and cites
 * the reference file:line it follows.  Where the reference's own FP32 bits are
 * not reproducible (it builds with --use_fast_math, CMakeLists.txt:17), this
 * file fixes a canonical operation order that the HIP kernels also use, so
 * that integer/topology results and FP32 node values agree bit for bit:
 *
 *   d2   = fmaf(dz, dz, fmaf(dy, dy, dx * dx))          (no eps)
 *   r2   = d2 + eps2
 *   inv  = 1 / sqrt(r2)                                  (GPU: v_rsq_f32)
 *   f    = ((m * inv) * inv) * inv
 *   a   += d * f   as fmaf(d, f, a)
 *   open = !(size2 < theta2 * d2),  size2 = size * size  (forces_bh.cu:356-359)
 *   COM  = FP64 combine of the two FP32 children, rounded once to FP32
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_spread21 / orc_morton : brute-force bit interleave over all 2^21
 *       inputs, closed forms, key30 == key63 >> 33, and the lbvh_check.py
 *       duplicate-key assertion (>= 23 duplicates on its own input).
 *   orc_sort_pairs            : numpy stable argsort.
 *   orc_karras                : an independent top-down highest-differing-bit
 *       split (orc_radix_topdown) + lbvh_check.py:103-118 topology checks.
 *   orc_com                   : FP64 recursive sums (numpy), root mass == total.
 *   orc_traverse_*            : theta=0 equals FP64 all-pairs to 1e-12 in the
 *       FP64 build of the walk and touches exactly n leaves
 *       (lbvh_check.py:250-264); theta=0.5 median error < 0.05 and
 *       interactions < n/2 (lbvh_check.py:266-279); at theta 0.3/0.5/0.8 the
 *       per-target and the 1/16/32/64-lane group walks' interaction and
 *       opening counts equal a Python restatement of lbvh_check.py:151-209
 *       (traverse, traverse_warp) on <= 3,024 points, walks with an opening
 *       test within 1e-6 of its threshold excepted
 *       (tests/test_walk_restatement.py).
 *   orc_allpairs_*            : two-body and three-body closed forms, Newton's
 *       third law (sum of m*a == 0), FP64 vs FP32 agreement.
 *   orc_energy_f64            : closed-form two-body energy; the committed
 *       reference trace benchmarks/energy_gpu_naive.csv (E0 = -0.1178396).
 *   orc_kdk_step              : closed-form constant-field step; energy drift
 *       over the 12k reference encounter.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_STACK 256

/* ---------------------------------------------------------------- Morton */

/* forces_bh.cu:192-200 (and lbvh_check.py:35-42): spread the low 21 bits of v
 * so bit k lands on bit 3k. */
uint64_t orc_spread21(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

/* Global AABB of the FP32 positions (forces_bh.cu:125-188); min/max are exact,
 * the quantisation constants are formed in FP64 as lbvh_check.py:45-48 does
 * for FP64 input: lo, ext = max_axis(hi - lo), inv = 1/ext (0 if ext == 0). */
void orc_quant_params(const float* pos4, int n, double lo[3], double* inv) {
  float l[3] = {INFINITY, INFINITY, INFINITY};
  float h[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      float v = pos4[4 * i + a];
      if (v < l[a]) l[a] = v;
      if (v > h[a]) h[a] = v;
    }
  double ext = 0.0;
  for (int a = 0; a < 3; ++a) {
    lo[a] = (double)l[a];
    double e = (double)h[a] - (double)l[a];
    if (e > ext) ext = e;
  }
  *inv = ext > 0.0 ? 1.0 / ext : 0.0;
}

/* One axis: q = clamp(trunc(((p - lo) * inv) * 2^21), 0, 2^21 - 1)
 * (lbvh_check.py:51-52; forces_bh.cu:212-217 does the same in FP32). */
static inline uint32_t quant21(float p, double lo, double inv) {
  double v = ((double)p - lo) * inv * 2097152.0;
  if (!(v > 0.0)) v = 0.0; /* also maps NaN to 0 */
  if (v > 2097151.0) v = 2097151.0;
  return (uint32_t)v;
}

/* Morton keys: 63-bit (21 bits/axis, forces_bh.cu:218) or 30-bit, which is
 * key63 >> 33 (10 bits/axis).  x takes bit 3k+2, y 3k+1, z 3k. */
void orc_morton(const float* pos4, int n, int key_bits, uint64_t* keys) {
  double lo[3], inv;
  if (n <= 0) return;
  orc_quant_params(pos4, n, lo, &inv);
  for (int i = 0; i < n; ++i) {
    uint32_t qx = quant21(pos4[4 * i + 0], lo[0], inv);
    uint32_t qy = quant21(pos4[4 * i + 1], lo[1], inv);
    uint32_t qz = quant21(pos4[4 * i + 2], lo[2], inv);
    uint64_t k = (orc_spread21(qx) << 2) | (orc_spread21(qy) << 1) | orc_spread21(qz);
    keys[i] = key_bits == 30 ? (k >> 33) : k;
  }
}

/* ------------------------------------------------------------------ sort */

typedef struct {
  uint64_t k;
  uint32_t i;
} orc_kv;

static void merge_sort_kv(orc_kv* a, orc_kv* tmp, int n) {
  /* bottom-up stable merge sort */
  for (int w = 1; w < n; w *= 2) {
    for (int lo = 0; lo < n; lo += 2 * w) {
      int mid = lo + w < n ? lo + w : n;
      int hi = lo + 2 * w < n ? lo + 2 * w : n;
      int p = lo, q = mid, o = lo;
      while (p < mid && q < hi) tmp[o++] = (a[q].k < a[p].k) ? a[q++] : a[p++];
      while (p < mid) tmp[o++] = a[p++];
      while (q < hi) tmp[o++] = a[q++];
    }
    memcpy(a, tmp, sizeof(orc_kv) * (size_t)n);
  }
}

/* Stable sort of (key, original index) pairs, as CUB SortPairs does with
 * vals[i] = i (forces_bh.cu:219, 490-497; lbvh_check.py:232 uses a stable
 * Python sort).  Ties keep ascending original index. */
int orc_sort_pairs(const uint64_t* keys, int n, uint64_t* keys_sorted, uint32_t* perm) {
  if (n <= 0) return 0;
  orc_kv* a = (orc_kv*)malloc(sizeof(orc_kv) * (size_t)n);
  orc_kv* t = (orc_kv*)malloc(sizeof(orc_kv) * (size_t)n);
  if (!a || !t) { free(a); free(t); return 1; }
  for (int i = 0; i < n; ++i) { a[i].k = keys[i]; a[i].i = (uint32_t)i; }
  merge_sort_kv(a, t, n);
  for (int i = 0; i < n; ++i) { keys_sorted[i] = a[i].k; perm[i] = a[i].i; }
  free(a);
  free(t);
  return 0;
}

/* ---------------------------------------------------------------- Karras */

static inline int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }
static inline int clz32(uint32_t x) { return x ? __builtin_clz(x) : 32; }

/* forces_bh.cu:234-240 / lbvh_check.py:61-67.  For 30-bit keys the build uses
 * a 32-bit key word, so delta is 32 smaller than the 64-bit form on both the
 * key and the tie branch; every comparison in the build is shift invariant,
 * so the topology is the same. */
static inline int delta_k(const uint64_t* k, int n, int bits, int i, int j) {
  if (j < 0 || j >= n) return -1;
  uint64_t x = k[i] ^ k[j];
  int wbits = bits == 30 ? 32 : 64;
  if (x) return wbits == 32 ? clz32((uint32_t)x) : clz64(x);
  return wbits + clz32((uint32_t)(i ^ j));
}

/* Karras 2012 radix tree, forces_bh.cu:242-277 / lbvh_check.py:57-100.
 * Internal nodes 0..n-2 (root 0), leaf i is id n-1+i.  children is 2*(n-1)
 * ints (left, right); parent is 2n-1 ints, parent[0] = -1. */
void orc_karras(const uint64_t* keys, int n, int key_bits, int32_t* children, int32_t* parent) {
  if (n < 2) { if (n == 1) parent[0] = -1; return; }
  for (int i = 0; i < 2 * n - 1; ++i) parent[i] = -1;
  for (int i = 0; i < n - 1; ++i) {
    int d = (delta_k(keys, n, key_bits, i, i + 1) - delta_k(keys, n, key_bits, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta_k(keys, n, key_bits, i, i - d);
    int lmax = 2;
    while (delta_k(keys, n, key_bits, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t > 0; t >>= 1)
      if (delta_k(keys, n, key_bits, i, i + (l + t) * d) > dmin) l += t;
    int j = i + l * d;
    int first = i < j ? i : j, last = i < j ? j : i;
    int common = delta_k(keys, n, key_bits, first, last);
    int split = first, step = last - first;
    do {
      step = (step + 1) >> 1;
      if (split + step < last && delta_k(keys, n, key_bits, first, split + step) > common) split += step;
    } while (step > 1);
    int left = (split == first) ? (n - 1 + split) : split;
    int right = (split + 1 == last) ? (n - 1 + split + 1) : split + 1;
    children[2 * i] = left;
    children[2 * i + 1] = right;
    parent[left] = i;
    parent[right] = i;
  }
}

/* Independent construction used only to pin orc_karras: split the sorted
 * range [lo, hi] top-down at the last position q whose extended key (key bits
 * then the 32 index bits) shares a longer prefix with keys[lo] than the whole
 * range does (found by a linear scan, not Karras's binary search).  Node ids
 * follow forces_bh.cu:272-273: a left-child range [lo, s] is internal node s,
 * a right-child range [s+1, hi] is internal node s+1, the root is node 0, and
 * one-element ranges are leaves n-1+index.  Returns 0 on success. */
typedef struct { int lo, hi, id; } orc_range;

static int ext_delta(const uint64_t* k, int n, int bits, int i, int j) {
  return delta_k(k, n, bits, i, j);
}

int orc_radix_topdown(const uint64_t* keys, int n, int key_bits, int32_t* children, int32_t* parent) {
  if (n < 2) { if (n == 1) parent[0] = -1; return 0; }
  orc_range* st = (orc_range*)malloc(sizeof(orc_range) * (size_t)(2 * n));
  if (!st) return 1;
  for (int i = 0; i < 2 * n - 1; ++i) parent[i] = -1;
  int sp = 0;
  st[sp++] = (orc_range){0, n - 1, 0};
  while (sp > 0) {
    orc_range r = st[--sp];
    int common = ext_delta(keys, n, key_bits, r.lo, r.hi);
    /* the split is the last s in [lo, hi-1] with delta(lo, s) > common;
     * found here by a linear scan (independent of the binary search) */
    int s = r.lo;
    for (int q = r.lo + 1; q < r.hi; ++q)
      if (ext_delta(keys, n, key_bits, r.lo, q) > common) s = q;
    int left_id, right_id;
    if (s == r.lo) left_id = n - 1 + s;
    else { left_id = s; st[sp++] = (orc_range){r.lo, s, s}; }
    if (s + 1 == r.hi) right_id = n - 1 + s + 1;
    else { right_id = s + 1; st[sp++] = (orc_range){s + 1, r.hi, s + 1}; }
    children[2 * r.id] = left_id;
    children[2 * r.id + 1] = right_id;
    parent[left_id] = r.id;
    parent[right_id] = r.id;
  }
  free(st);
  return 0;
}

/* ------------------------------------------------------------------- COM */

/* Bottom-up centres of mass and AABBs, forces_bh.cu:281-331 and
 * lbvh_check.py:121-148.  The combine is done in FP64 from the FP32 child
 * values and rounded once:
 *   m = ml + mr;  c = (cl*ml + cr*mr) / m;  box = union;  size = longest edge.
 * com4 is (x, y, z, m) per internal node, size2 = size*size in FP32.
 * Processed in reverse id order of a post-order walk so children are ready. */
int orc_com(const float* pos_sorted4, int n, const int32_t* children, float* com4,
            float* box_lo4, float* box_hi4, float* size2) {
  if (n < 2) return 0;
  int* order = (int*)malloc(sizeof(int) * (size_t)(n - 1));
  int* st = (int*)malloc(sizeof(int) * (size_t)(2 * n));
  if (!order || !st) { free(order); free(st); return 1; }
  int no = 0, sp = 0;
  st[sp++] = 0;
  while (sp > 0) {
    int id = st[--sp];
    if (id < n - 1) {
      order[no++] = id;
      st[sp++] = children[2 * id];
      st[sp++] = children[2 * id + 1];
    }
  }
  for (int k = no - 1; k >= 0; --k) {
    int id = order[k];
    float c[2][4], l[2][3], h[2][3];
    for (int s = 0; s < 2; ++s) {
      int ch = children[2 * id + s];
      if (ch >= n - 1) {
        const float* p = pos_sorted4 + 4 * (size_t)(ch - (n - 1));
        for (int a = 0; a < 4; ++a) c[s][a] = p[a];
        for (int a = 0; a < 3; ++a) l[s][a] = h[s][a] = p[a];
      } else {
        for (int a = 0; a < 4; ++a) c[s][a] = com4[4 * (size_t)ch + a];
        for (int a = 0; a < 3; ++a) {
          l[s][a] = box_lo4[4 * (size_t)ch + a];
          h[s][a] = box_hi4[4 * (size_t)ch + a];
        }
      }
    }
    double ml = (double)c[0][3], mr = (double)c[1][3];
    double m = ml + mr;
    for (int a = 0; a < 3; ++a) {
      double num = (double)c[0][a] * ml + (double)c[1][a] * mr;
      com4[4 * (size_t)id + a] = (float)(num / m);
    }
    com4[4 * (size_t)id + 3] = (float)m;
    float lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = fminf(l[0][a], l[1][a]);
      hi[a] = fmaxf(h[0][a], h[1][a]);
      box_lo4[4 * (size_t)id + a] = lo[a];
      box_hi4[4 * (size_t)id + a] = hi[a];
    }
    box_lo4[4 * (size_t)id + 3] = 0.f;
    box_hi4[4 * (size_t)id + 3] = 0.f;
    float s = fmaxf(hi[0] - lo[0], fmaxf(hi[1] - lo[1], hi[2] - lo[2]));
    size2[id] = s * s;
  }
  free(order);
  free(st);
  return 0;
}

/* ------------------------------------------------------------- force law */

static inline float orc_d2(float dx, float dy, float dz) {
  return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

/* body_body, forces.cu:17-31, in the canonical order above. */
static inline void orc_interact(float bx, float by, float bz, const float* src,
                                float eps2, float a[3]) {
  float dx = src[0] - bx, dy = src[1] - by, dz = src[2] - bz;
  float r2 = orc_d2(dx, dy, dz) + eps2;
  float inv = 1.0f / sqrtf(r2);
  float f = src[3] * inv * inv * inv;
  a[0] = fmaf(dx, f, a[0]);
  a[1] = fmaf(dy, f, a[1]);
  a[2] = fmaf(dz, f, a[2]);
}

/* Naive all-pairs in FP32 with the reference kernel's j order: sources j
 * ascending, self term exactly zero (forces.cu:33-52).  acc4 gets G*a. */
void orc_allpairs_f32(const float* pos4, int n, float eps, float G, float* acc4) {
  float eps2 = eps * eps;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    float a[3] = {0.f, 0.f, 0.f};
    float bx = pos4[4 * i], by = pos4[4 * i + 1], bz = pos4[4 * i + 2];
    for (int j = 0; j < n; ++j) orc_interact(bx, by, bz, pos4 + 4 * (size_t)j, eps2, a);
    acc4[4 * i + 0] = G * a[0];
    acc4[4 * i + 1] = G * a[1];
    acc4[4 * i + 2] = G * a[2];
    acc4[4 * i + 3] = 0.f;
  }
}

/* FP64 all-pairs on selected targets (the FP64 oracle of the accuracy gates;
 * law of forces.cu:12-15 / reference_nbody.py:18-19).  targets may be NULL
 * for all particles; acc3 is 3*nt doubles. */
void orc_allpairs_f64(const float* pos4, int n, double eps, double G, const int32_t* targets,
                      int nt, double* acc3) {
  double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < nt; ++k) {
    int i = targets ? targets[k] : k;
    double bx = pos4[4 * i], by = pos4[4 * i + 1], bz = pos4[4 * i + 2];
    double ax = 0, ay = 0, az = 0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = pos4[4 * j] - bx, dy = pos4[4 * j + 1] - by, dz = pos4[4 * j + 2] - bz;
      double r2 = dx * dx + dy * dy + dz * dz + eps2;
      double f = pos4[4 * j + 3] / (r2 * sqrt(r2));
      ax += dx * f;
      ay += dy * f;
      az += dz * f;
    }
    acc3[3 * k] = G * ax;
    acc3[3 * k + 1] = G * ay;
    acc3[3 * k + 2] = G * az;
  }
}

/* ------------------------------------------------------------- traversal */

/* Per-target stack walk, forces_bh.cu:335-377 (lbvh_check.py:151-177).
 * Targets are sorted indices t in [t0, t1).  Writes G*a in *sorted* order to
 * acc4_sorted[t] (the caller scatters by perm), and per-target interaction
 * and opening-test counts.  Returns the maximum stack depth, or -1 on
 * overflow. */
int orc_traverse_thread(const float* pos_sorted4, int n, const int32_t* children,
                        const float* com4, const float* size2, float theta, float eps,
                        float G, int t0, int t1, float* acc4_sorted, int32_t* inter,
                        int32_t* opens) {
  float theta2 = theta * theta, eps2 = eps * eps;
  int worst = 0, overflow = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : worst) reduction(| : overflow)
  for (int t = t0; t < t1; ++t) {
    int st[ORC_STACK];
    int sp = 0, ni = 0, no = 0;
    float bx = pos_sorted4[4 * t], by = pos_sorted4[4 * t + 1], bz = pos_sorted4[4 * t + 2];
    float a[3] = {0.f, 0.f, 0.f};
    st[sp++] = 0;
    while (sp > 0) {
      if (sp > worst) worst = sp;
      int id = st[--sp];
      const float* src;
      if (n == 1 || id >= n - 1) {
        src = pos_sorted4 + 4 * (size_t)(id - (n - 1));
      } else {
        const float* c = com4 + 4 * (size_t)id;
        float d2 = orc_d2(c[0] - bx, c[1] - by, c[2] - bz);
        ++no;
        if (size2[id] < theta2 * d2) {
          src = c;
        } else {
          if (sp + 2 > ORC_STACK) { overflow = 1; sp = 0; break; }
          st[sp++] = children[2 * id];
          st[sp++] = children[2 * id + 1];
          continue;
        }
      }
      orc_interact(bx, by, bz, src, eps2, a);
      ++ni;
    }
    if (acc4_sorted) {
      acc4_sorted[4 * t] = G * a[0];
      acc4_sorted[4 * t + 1] = G * a[1];
      acc4_sorted[4 * t + 2] = G * a[2];
      acc4_sorted[4 * t + 3] = 0.f;
    }
    if (inter) inter[t] = ni;
    if (opens) opens[t] = no;
  }
  return overflow ? -1 : worst;
}

/* W-lane group walk: forces_bh.cu:395-447 with the 32-lane warp replaced by a
 * group of `width` consecutive sorted targets (lbvh_check.py:180-209).  One
 * shared stack; a node opens when any live lane's test opens it; otherwise
 * every lane takes it (or the leaf) as a source.  Per-group interaction and
 * opening-test counts (one count per popped node, not per lane).  Returns max
 * depth or -1 on overflow. */
static int group_walk(const float* pos_sorted4, int n, const int32_t* children,
                      const float* com4, const float* size2, float theta, float eps, float G,
                      int width, float* acc4_sorted, double* acc3_sorted_f64,
                      int32_t* group_inter, int32_t* group_opens);

int orc_traverse_group(const float* pos_sorted4, int n, const int32_t* children,
                       const float* com4, const float* size2, float theta, float eps, float G,
                       int width, float* acc4_sorted, int32_t* group_inter,
                       int32_t* group_opens) {
  return group_walk(pos_sorted4, n, children, com4, size2, theta, eps, G, width, acc4_sorted,
                    NULL, group_inter, group_opens);
}

/* The same group walk -- identical FP32 open/accept decisions, so identical
 * interaction lists -- with each lane's list summed in FP64 from the FP32
 * source and target values (d, r^2 + eps^2 with the walk's FP32 eps^2, the
 * force factor and the sum all in FP64).  Against it, a force module's FP32
 * result shows its arithmetic error alone (rounding of the per-interaction
 * terms and of their accumulation), separated from the multipole error of
 * the list itself.  acc3_sorted is 3 doubles per sorted target. */
int orc_traverse_group_f64sum(const float* pos_sorted4, int n, const int32_t* children,
                              const float* com4, const float* size2, float theta, float eps,
                              double G, int width, double* acc3_sorted, int32_t* group_inter) {
  return group_walk(pos_sorted4, n, children, com4, size2, theta, eps, (float)G, width, NULL,
                    acc3_sorted, group_inter, NULL);
}

static int group_walk(const float* pos_sorted4, int n, const int32_t* children,
                      const float* com4, const float* size2, float theta, float eps, float G,
                      int width, float* acc4_sorted, double* acc3_sorted_f64,
                      int32_t* group_inter, int32_t* group_opens) {
  float theta2 = theta * theta, eps2 = eps * eps;
  int ngroups = (n + width - 1) / width;
  int worst = 0, overflow = 0;
#pragma omp parallel for schedule(dynamic, 4) reduction(max : worst) reduction(| : overflow)
  for (int g = 0; g < ngroups; ++g) {
    int base = g * width;
    int nl = n - base < width ? n - base : width;
    float* a = (float*)calloc((size_t)nl * 3, sizeof(float));
    double* a64 = (double*)calloc((size_t)nl * 3, sizeof(double));
    int st[ORC_STACK];
    int sp = 0, ni = 0, no = 0;
    st[sp++] = 0;
    while (sp > 0) {
      if (sp > worst) worst = sp;
      int id = st[--sp];
      const float* src;
      if (n == 1 || id >= n - 1) {
        src = pos_sorted4 + 4 * (size_t)(id - (n - 1));
      } else {
        const float* c = com4 + 4 * (size_t)id;
        int open = 0;
        for (int l = 0; l < nl; ++l) {
          const float* b = pos_sorted4 + 4 * (size_t)(base + l);
          float d2 = orc_d2(c[0] - b[0], c[1] - b[1], c[2] - b[2]);
          if (!(size2[id] < theta2 * d2)) open = 1;
        }
        ++no;
        if (open) {
          if (sp + 2 > ORC_STACK) { overflow = 1; break; }
          st[sp++] = children[2 * id];
          st[sp++] = children[2 * id + 1];
          continue;
        }
        src = c;
      }
      for (int l = 0; l < nl; ++l) {
        const float* b = pos_sorted4 + 4 * (size_t)(base + l);
        if (acc4_sorted) orc_interact(b[0], b[1], b[2], src, eps2, a + 3 * l);
        if (acc3_sorted_f64) {
          const double dx = (double)src[0] - b[0], dy = (double)src[1] - b[1],
                       dz = (double)src[2] - b[2];
          const double r2 = dx * dx + dy * dy + dz * dz + (double)eps2;
          const double f = (double)src[3] / (r2 * sqrt(r2));
          a64[3 * l] += dx * f;
          a64[3 * l + 1] += dy * f;
          a64[3 * l + 2] += dz * f;
        }
      }
      ++ni;
    }
    if (acc4_sorted)
      for (int l = 0; l < nl; ++l) {
        acc4_sorted[4 * (size_t)(base + l)] = G * a[3 * l];
        acc4_sorted[4 * (size_t)(base + l) + 1] = G * a[3 * l + 1];
        acc4_sorted[4 * (size_t)(base + l) + 2] = G * a[3 * l + 2];
        acc4_sorted[4 * (size_t)(base + l) + 3] = 0.f;
      }
    if (acc3_sorted_f64)
      for (int l = 0; l < nl * 3; ++l)
        acc3_sorted_f64[3 * (size_t)base + l] = (double)G * a64[l];
    if (group_inter) group_inter[g] = ni;
    if (group_opens) group_opens[g] = no;
    free(a);
    free(a64);
  }
  return overflow ? -1 : worst;
}

/* FP64 twin of the thread walk on FP64 node data (used only to pin the walk
 * logic: theta=0 must equal FP64 all-pairs to 1e-12, lbvh_check.py:262).
 * com4d is (x,y,z,m) FP64 per internal node, size2d FP64. */
int orc_traverse_thread_f64(const double* pos_sorted4, int n, const int32_t* children,
                            const double* com4d, const double* size2d, double theta,
                            double eps, int t0, int t1, double* acc3, int32_t* inter) {
  double theta2 = theta * theta, eps2 = eps * eps;
  int worst = 0;
  for (int t = t0; t < t1; ++t) {
    int st[ORC_STACK];
    int sp = 0, ni = 0;
    const double* b = pos_sorted4 + 4 * (size_t)t;
    double ax = 0, ay = 0, az = 0;
    st[sp++] = 0;
    while (sp > 0) {
      if (sp > worst) worst = sp;
      int id = st[--sp];
      const double* src;
      if (n == 1 || id >= n - 1) {
        src = pos_sorted4 + 4 * (size_t)(id - (n - 1));
      } else {
        const double* c = com4d + 4 * (size_t)id;
        double dx = c[0] - b[0], dy = c[1] - b[1], dz = c[2] - b[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        if (size2d[id] < theta2 * d2) {
          src = c;
        } else {
          if (sp + 2 > ORC_STACK) return -1;
          st[sp++] = children[2 * id];
          st[sp++] = children[2 * id + 1];
          continue;
        }
      }
      double dx = src[0] - b[0], dy = src[1] - b[1], dz = src[2] - b[2];
      double r2 = dx * dx + dy * dy + dz * dz + eps2;
      double f = src[3] / (r2 * sqrt(r2));
      ax += dx * f;
      ay += dy * f;
      az += dz * f;
      ++ni;
    }
    acc3[3 * (size_t)(t - t0)] = ax;
    acc3[3 * (size_t)(t - t0) + 1] = ay;
    acc3[3 * (size_t)(t - t0) + 2] = az;
    if (inter) inter[t - t0] = ni;
  }
  return worst;
}

/* -------------------------------------------------------- energy and KDK */

/* Total energy in FP64, integrator.cu:39-110 / reference_nbody.py:64-81:
 *   KE = sum 1/2 m |v|^2,  PE = -1/2 G sum_{i != j} m_i m_j / sqrt(r^2 + eps^2).
 * out = {KE + PE, KE, PE}. */
void orc_energy_f64(const float* pos4, const float* vel4, int n, double eps, double G,
                    double out[3]) {
  double eps2 = eps * eps, ke = 0.0, pe = 0.0;
  for (int i = 0; i < n; ++i) {
    double vx = vel4[4 * i], vy = vel4[4 * i + 1], vz = vel4[4 * i + 2];
    ke += 0.5 * (double)pos4[4 * i + 3] * (vx * vx + vy * vy + vz * vz);
  }
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : pe)
  for (int i = 0; i < n; ++i) {
    double bx = pos4[4 * i], by = pos4[4 * i + 1], bz = pos4[4 * i + 2];
    double phi = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      double dx = pos4[4 * j] - bx, dy = pos4[4 * j + 1] - by, dz = pos4[4 * j + 2] - bz;
      phi += (double)pos4[4 * j + 3] / sqrt(dx * dx + dy * dy + dz * dz + eps2);
    }
    pe += (double)pos4[4 * i + 3] * phi;
  }
  pe *= -0.5 * G;
  out[0] = ke + pe;
  out[1] = ke;
  out[2] = pe;
}

/* Kick and drift halves of leapfrog_step, integrator.cu:5-37:
 *   v += (dt/2) a;  x += dt v    (mass in pos.w untouched). */
void orc_kick(float* vel4, const float* acc4, int n, float half_dt) {
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) vel4[4 * i + a] = fmaf(half_dt, acc4[4 * i + a], vel4[4 * i + a]);
}

void orc_drift(float* pos4, const float* vel4, int n, float dt) {
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) pos4[4 * i + a] = fmaf(dt, vel4[4 * i + a], pos4[4 * i + a]);
}

/* One naive KDK step (integrator.cu:27-37 with forces_naive):
 * acc4 holds a(x) on entry and a(x_new) on exit. */
void orc_kdk_step_naive(float* pos4, float* vel4, float* acc4, int n, float dt, float eps,
                        float G) {
  float half = 0.5f * dt;
  orc_kick(vel4, acc4, n, half);
  orc_drift(pos4, vel4, n, dt);
  orc_allpairs_f32(pos4, n, eps, G, acc4);
  orc_kick(vel4, acc4, n, half);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
