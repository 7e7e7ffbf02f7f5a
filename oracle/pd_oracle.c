/*
 * pd_oracle.c -- CPU restatement of the PeriCUDA implicit peridynamics
 * step.  TEST INFRASTRUCTURE ONLY (see pd_oracle.h): the parity checker and
 * the timed CPU baseline, never part of the product path.
 *
 * Build: gcc -O2 -fopenmp -ffp-contract=off -shared -fPIC (oracle/Makefile).
 * OpenMP use is deterministic: every parallel loop writes only its own rows
 * and every reduction runs in the fixed order of orc_dot below.
 */
#include "pd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* core/definitions.h constants used on the path */
#define ORC_SMALL 1.0e-10   /* definitions.h:61 */
#define ORC_ALPHA_PR 0.25   /* definitions.h:109 */

/* ======================================================================
 * glibc rand(): TYPE_3 additive feedback generator, the algorithm of
 * glibc random_r.c (srandom_r + random_r with the default 128-byte state),
 * restated from its published description:
 *   r_0 = seed (0 -> 1); r_i = 16807 r_{i-1} mod (2^31-1), i = 1..30
 *   (Schrage's method); r_i = r_{i-31}, i = 31..33;
 *   r_i = r_{i-31} + r_{i-3} mod 2^32, i >= 34;
 *   the k-th output (k = 0,1,...) is r_{k+344} >> 1.
 * Pinned against the C library's own rand() in tests/test_oracle_rand.py.
 * ====================================================================== */
static inline uint32_t ring_get(const orc_rand* st, int64_t i) {
  return st->ring[i % 34];
}

void orc_rand_seed(orc_rand* st, uint32_t seed) {
  int32_t r = (int32_t)(seed == 0 ? 1u : seed);
  st->ring[0] = (uint32_t)r;
  for (int i = 1; i < 31; ++i) {
    int64_t hi = r / 127773;
    int64_t lo = r % 127773;
    int64_t word = 16807 * lo - 2836 * hi;
    if (word < 0) word += 2147483647;
    r = (int32_t)word;
    st->ring[i] = (uint32_t)r;
  }
  for (int i = 31; i < 34; ++i) st->ring[i] = st->ring[i - 31];
  st->i = 34;
  for (int k = 34; k < 344; ++k) (void)orc_rand_next(st);
}

int32_t orc_rand_next(orc_rand* st) {
  int64_t i = st->i;
  uint32_t v = ring_get(st, i - 31) + ring_get(st, i - 3);
  st->ring[i % 34] = v;
  st->i = i + 1;
  return (int32_t)(v >> 1);
}

/* core/simulator.cu:802-805 frand(): rand() / (real_t) RAND_MAX */
static inline double orc_frand(orc_rand* st) {
  return (double)orc_rand_next(st) / (double)2147483647;
}

/* core/simulator.cu:807-823 getstretch(): s0 arrives as the float
 * RunningParameters::pd_stretch_limit_s0 (core/parameters.h:123). */
double orc_getstretch(orc_rand* st, float s0f) {
  double s0 = (double)s0f;
  double sum = 0.0;
  for (int i = 0; i < 12; ++i) sum += orc_frand(st);
  sum -= 6.0;
  return ((sum * 0.01 + 1.0) * s0);
}

/* core/simulator.cu:855-864 */
void orc_stretch_init(int32_t n, float s0, uint32_t seed, double* s00) {
  orc_rand st;
  orc_rand_seed(&st, seed);
  for (int32_t i = 0; i < n; ++i) s00[i] = orc_getstretch(&st, s0);
}

/* ======================================================================
 * Bond lists
 * ====================================================================== */

/* core/kdtree.h:66-79 dist(): sqrt(((0 + dx^2) + dy^2) + dz^2), dx = a-b */
static inline double kd_dist(const double* a, const double* b) {
  double howfar = 0.0;
  for (int j = 0; j < 3; ++j) howfar += (a[j] - b[j]) * (a[j] - b[j]);
  return sqrt(howfar);
}

typedef struct kd_point {
  double c[3];
  int32_t index;
} kd_point;

typedef struct kd_node {
  struct kd_node* left;
  struct kd_node* right;
  int axis;
  double split;
  int isleaf;
  kd_point* point;
  int count;
  double box_m[3];
  double box_M[3];
} kd_node;

/* core/kdtree.h:82-115 dist_pt_to_box */
static double kd_dist_pt_to_box(const kd_point* t, const double* m,
                                const double* M) {
  double hit[3];
  for (int j = 0; j < 3; ++j) {
    hit[j] = t->c[j];
    if (t->c[j] < m[j]) hit[j] = m[j];
    else if (t->c[j] > M[j]) hit[j] = M[j];
  }
  return kd_dist(hit, t->c);
}

/* core/kdtree.h:168-227 get_median: Hoare-partition quickselect of the
 * k-th smallest along `axis`, k = size/2 - 1 (even) or size/2 (odd). */
static double kd_get_median(kd_point* point, int size, int axis) {
  int k = (size % 2 == 0) ? size / 2 - 1 : size / 2;
  int left = 0, right = size - 1;
  while (left < right) {
    double key = point[k].c[axis];
    int i = left, j = right;
    do {
      while (point[i].c[axis] < key) i++;
      while (point[j].c[axis] > key) j--;
      if (i <= j) {
        kd_point hold = point[i];
        point[i] = point[j];
        point[j] = hold;
        i++;
        j--;
      }
    } while (i <= j);
    if (j < k) left = i;
    if (k < i) right = j;
  }
  return point[k].c[axis];
}

static kd_node* kd_new_node(kd_point* p, int count, const double* m,
                            const double* M) {
  kd_node* nd = (kd_node*)calloc(1, sizeof(kd_node));
  nd->axis = -1;
  nd->isleaf = 1;
  nd->point = p;
  nd->count = count;
  memcpy(nd->box_m, m, sizeof(double) * 3);
  memcpy(nd->box_M, M, sizeof(double) * 3);
  return nd;
}

/* core/kdtree.h:229-288 buildtree: split the longest box axis at the
 * median; left keeps ceil(count/2) points, right floor(count/2). */
static void kd_build(kd_node* nd, int max_items) {
  if (nd->isleaf) {
    int lc = nd->count;
    if (lc <= max_items) return;
    double maxlength = 0, up[3], dn[3];
    int longaxis = 0;
    for (int j = 0; j < 3; ++j) {
      up[j] = nd->box_M[j];
      dn[j] = nd->box_m[j];
      double length = up[j] - dn[j];
      if (length > maxlength) {
        longaxis = j;
        maxlength = length;
      }
    }
    double median = kd_get_median(nd->point, lc, longaxis);
    nd->split = median;
    nd->axis = longaxis;
    nd->isleaf = 0;
    int leftcount = (lc % 2 == 0) ? lc / 2 : lc / 2 + 1;
    up[longaxis] = dn[longaxis] = median;
    nd->left = kd_new_node(nd->point, leftcount, nd->box_m, up);
    nd->right = kd_new_node(nd->point + leftcount, lc / 2, dn, nd->box_M);
  }
  kd_build(nd->left, max_items);
  kd_build(nd->right, max_items);
}

static void kd_free(kd_node* nd) {
  if (!nd) return;
  kd_free(nd->left);
  kd_free(nd->right);
  free(nd);
}

typedef struct kd_out {
  int32_t* counts;
  int32_t* lists;
  int32_t max_bonds;
  double horizon;
  int64_t dropped;
} kd_out;

/* core/kdtree.h:304-353 find_neighbors + add_bond */
static void kd_find(const kd_point* t, const kd_node* nd, kd_out* o) {
  if (kd_dist_pt_to_box(t, nd->box_m, nd->box_M) > o->horizon) return;
  if (nd->isleaf) {
    int32_t j = t->index;
    for (int i = 0; i < nd->count; ++i) {
      double r = kd_dist(t->c, nd->point[i].c);
      if (r < o->horizon) {
        int32_t k = nd->point[i].index;
        if (j != k) {
          if (o->counts[j] >= o->max_bonds) {
            o->dropped++;
          } else {
            o->lists[(int64_t)j * o->max_bonds + o->counts[j]] = k;
            o->counts[j]++;
          }
        }
      }
    }
    return;
  }
  kd_find(t, nd->left, o);
  kd_find(t, nd->right, o);
}

static int cmp_i32(const void* a, const void* b) {
  int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
  return (x > y) - (x < y);
}

/* core/simulator.cu:825-943 initPeridynamicsBonds (bond part) */
int64_t orc_bonds_kdtree(int32_t n, const double* x4, double horizon,
                         int32_t max_bonds, int32_t leaf_size,
                         int32_t* counts, int32_t* lists) {
  kd_point* pts = (kd_point*)malloc(sizeof(kd_point) * (size_t)(n > 0 ? n : 1));
  for (int32_t i = 0; i < n; ++i) {
    pts[i].c[0] = x4[4 * (int64_t)i + 0];
    pts[i].c[1] = x4[4 * (int64_t)i + 1];
    pts[i].c[2] = x4[4 * (int64_t)i + 2];
    pts[i].index = i;
    counts[i] = 0;
  }
  /* query points keep the original order (core/simulator.cu:862, 875) */
  kd_point* query = (kd_point*)malloc(sizeof(kd_point) * (size_t)(n > 0 ? n : 1));
  memcpy(query, pts, sizeof(kd_point) * (size_t)n);
  const double tiny[3] = {-1000000.0, -1000000.0, -1000000.0};
  const double huge[3] = {1000000.0, 1000000.0, 1000000.0};
  kd_node* root = kd_new_node(pts, n, tiny, huge);
  kd_build(root, leaf_size);
  kd_out o = {counts, lists, max_bonds, horizon, 0};
  /* Rows are independent; each query writes only its own row. */
  int64_t dropped = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : dropped)
  for (int32_t i = 0; i < n; ++i) {
    kd_out oi = o;
    oi.dropped = 0;
    kd_find(&query[i], root, &oi);
    dropped += oi.dropped;
  }
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i)
    qsort(lists + (int64_t)i * max_bonds, (size_t)counts[i], sizeof(int32_t),
          cmp_i32);
  kd_free(root);
  free(pts);
  free(query);
  return dropped;
}

int64_t orc_bonds_brute(int32_t n, const double* x4, const int32_t* body,
                        const double* horizon, int32_t max_bonds,
                        int32_t* counts, int32_t* lists) {
  int overflow = 0;
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total) \
    reduction(| : overflow)
  for (int32_t p = 0; p < n; ++p) {
    const double* a = x4 + 4 * (int64_t)p;
    int bp = body ? body[p] : 0;
    double h = horizon[bp];
    int32_t c = 0;
    for (int32_t q = 0; q < n; ++q) {
      if (q == p) continue;
      if (body && body[q] != bp) continue;
      double r = kd_dist(a, x4 + 4 * (int64_t)q);
      if (r < h) {
        if (c < max_bonds) lists[(int64_t)p * max_bonds + c] = q;
        else overflow = 1;
        c++;
      }
    }
    counts[p] = c < max_bonds ? c : max_bonds;
    total += c;
  }
  return overflow ? -1 : total;
}

/* ======================================================================
 * Deterministic reduction (shared with the HIP kernels)
 *
 * Rows are cut into tiles of 256 consecutive particles; tile g is one
 * 256-thread workgroup on the GPU, thread t owning row 256 g + t.
 *   term_i   = (a.x*b.x + a.y*b.y) + a.z*b.z        (no FMA)
 *   per wave (64 lanes) a xor-butterfly: v_l <- v_l + v_{l^m},
 *            m = 32,16,8,4,2,1 (every lane ends with the same value)
 *   per tile (w0 + w1) + (w2 + w3) over its four waves -> partial[g]
 *   final    thread t sums partial[t], partial[t+256], ... left to right
 *            from 0.0, then the same butterfly and (w0+w1)+(w2+w3).
 * Absent rows (i >= n) contribute 0.0.
 * ====================================================================== */
#define ORC_TILE 256

static double butterfly64(double* v) {
  for (int m = 32; m >= 1; m >>= 1) {
    double t[64];
    for (int l = 0; l < 64; ++l) t[l] = v[l] + v[l ^ m];
    memcpy(v, t, sizeof(t));
  }
  return v[0];
}

static double tile_combine(double* v256) {
  double w0 = butterfly64(v256 + 0);
  double w1 = butterfly64(v256 + 64);
  double w2 = butterfly64(v256 + 128);
  double w3 = butterfly64(v256 + 192);
  return (w0 + w1) + (w2 + w3);
}

static double final_combine(int64_t g, const double* part) {
  double v[256];
  for (int t = 0; t < 256; ++t) {
    double acc = 0.0;
    for (int64_t k = t; k < g; k += 256) acc = acc + part[k];
    v[t] = acc;
  }
  return tile_combine(v);
}

static inline double term3(const double* a, const double* b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

double orc_dot(int32_t n, const double* a4, const double* b4) {
  int64_t g = (n + ORC_TILE - 1) / ORC_TILE;
  if (g == 0) return final_combine(0, NULL);
  double* part = (double*)malloc(sizeof(double) * (size_t)g);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < g; ++t) {
    double v[256];
    for (int l = 0; l < 256; ++l) {
      int64_t i = t * ORC_TILE + l;
      v[l] = (i < n) ? term3(a4 + 4 * i, b4 + 4 * i) : 0.0;
    }
    part[t] = tile_combine(v);
  }
  double r = final_combine(g, part);
  free(part);
  return r;
}

/* orc_dot's tree over precomputed per-row terms */
static double orc_tree(int32_t n, const double* terms) {
  int64_t g = (n + ORC_TILE - 1) / ORC_TILE;
  if (g == 0) return final_combine(0, NULL);
  double* part = (double*)malloc(sizeof(double) * (size_t)g);
#pragma omp parallel for schedule(static)
  for (int64_t t = 0; t < g; ++t) {
    double v[256];
    for (int l = 0; l < 256; ++l) {
      int64_t i = t * ORC_TILE + l;
      v[l] = (i < n) ? terms[i] : 0.0;
    }
    part[t] = tile_combine(v);
  }
  double r = final_combine(g, part);
  free(part);
  return r;
}

/* FP64 product of FP32 operands: every float is widened exactly */
static inline double term3f(const float* a, const float* b) {
  return ((double)a[0] * (double)b[0] + (double)a[1] * (double)b[1]) +
         (double)a[2] * (double)b[2];
}

double orc_dot_f32(int32_t n, const float* a4, const float* b4) {
  double* t = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) t[i] = term3f(a4 + 4 * (int64_t)i, b4 + 4 * (int64_t)i);
  const double r = orc_tree(n, t);
  free(t);
  return r;
}

/* dot over three vector pairs in one sweep (same tree for each) */
static void orc_dot3(int32_t n, const double* a0, const double* b0,
                     const double* a1, const double* b1, const double* a2,
                     const double* b2, double* out) {
  out[0] = orc_dot(n, a0, b0);
  out[1] = a1 ? orc_dot(n, a1, b1) : 0.0;
  out[2] = a2 ? orc_dot(n, a2, b2) : 0.0;
}

/* ======================================================================
 * Assembly: force, 3x3 Jacobian, breaking, and the integrator's fill
 * ====================================================================== */

/* helper_math.h:43-46 host fmaxf(float,float): a > b ? a : b */
static inline float orc_fmaxf(float a, float b) { return a > b ? a : b; }

void orc_assemble(orc_system* sys, const orc_step_params* prm,
                  orc_step_stats* st) {
  const int32_t n = sys->n;
  const double clock = prm->clock_scale;
  const double dt = prm->dt;
  /* implicit_euler.cuh:39-40 / newmark_beta.cuh:39-40 */
  double dt2 = dt * dt;
  if (prm->integrator == 1) dt2 = (dt * dt) / 4.0;
  int64_t broken = 0, live = 0;
  double gmin = INFINITY;
  int32_t gp = -1, gq = -1;

#pragma omp parallel
  {
    double lmin = INFINITY;
    int32_t lp = -1, lq = -1;
#pragma omp for schedule(static) reduction(+ : broken, live)
    for (int32_t p = 0; p < n; ++p) {
      if (sys->act[p] == ORC_INVISIBLE) continue; /* cg_solver.cuh:225-228 */
      const orc_material* mt = &sys->mat[sys->body[p]];
      const double* pos_p = sys->x + 4 * (int64_t)p;
      const double* opos_p = sys->x0 + 4 * (int64_t)p;
      double force[4] = {0.0, 0.0, 0.0, 0.0};
      const int64_t rs = sys->row_start[p];
      int32_t top = sys->top[p];
      const int32_t otop = top;
      const double stretch = sys->s[p];
      double smin = 0.0;
      double sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      int32_t bnd = 0;
      /* cg_solver.cuh:249-319 */
      while (bnd <= top) {
        const int32_t q = sys->col[rs + bnd];
        const double* xq = sys->x + 4 * (int64_t)q;
        double e[4] = {xq[0] - pos_p[0], xq[1] - pos_p[1], xq[2] - pos_p[2],
                       xq[3] - pos_p[3]};
        const double dij = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
        /* cg_solver.cuh:256 eij /= fmaxf(dij, SMALL): float rounding */
        const double fd = (double)orc_fmaxf((float)dij, (float)ORC_SMALL);
        e[0] /= fd;
        e[1] /= fd;
        e[2] /= fd;
        e[3] /= fd;
        const double* oq = sys->x0 + 4 * (int64_t)q;
        const double o[4] = {opos_p[0] - oq[0], opos_p[1] - oq[1],
                             opos_p[2] - oq[2], opos_p[3] - oq[3]};
        /* cutil_math_ext.h length(double4) = sqrt(dot4) */
        const double d0 =
            sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2] + o[3] * o[3]);
        const double drdivd = dij / d0 - 1.0;
        /* cg_solver.cuh:264 fmaxf(stretch, pd_stretch[q]): float rounding */
        const double thr =
            (double)orc_fmaxf((float)stretch, (float)sys->s[q]);
        const double gap = fabs(drdivd - thr);
        if (gap < lmin || (gap == lmin && p < lp)) {
          lmin = gap;
          lp = p;
          lq = q;
        }
        if (drdivd > thr) {
          /* cg_solver.cuh:266-275 swap-remove, walker unchanged */
          const int32_t last = sys->col[rs + top];
          sys->col[rs + bnd] = last;
          sys->col[rs + top] = q;
          --top;
          continue;
        }
        double vscale = 1.0;
        if (d0 > (mt->horizon - mt->radius))
          vscale = 0.5 + (mt->horizon - d0) / 2.0 / mt->radius;
        if (drdivd < smin) smin = drdivd;
        const double kk = drdivd * vscale;
        force[0] += kk * e[0];
        force[1] += kk * e[1];
        force[2] += kk * e[2];
        force[3] += kk * e[3];
        /* cg_solver.cuh:292-299 */
        const double ratio = d0 / dij;
        const double dg = 1.0 - ratio;
        const double sc = mt->c_bond / d0;
        double F[9];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) {
            const double xx = (e[i] * e[j]) * ratio;
            const double base = (i == j) ? dg : 0.0;
            F[3 * i + j] = (base + xx) * sc;
          }
        for (int k = 0; k < 9; ++k) sum[k] += F[k];
        /* stored block -J, then the integrator scales it by dt^2
         * (implicit_euler.cuh:50-57, newmark_beta.cuh:50-57) */
        double* out = sys->val + 9 * (rs + bnd);
        for (int k = 0; k < 9; ++k) out[k] = (F[k] * -1.0) * dt2;
        ++bnd;
      }
      /* diagonal: A_pp = K_pp dt^2 + m I (implicit_euler.cuh:42-48) */
      double* dgo = sys->diag + 9 * (int64_t)p;
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
          const double mm = (i == j) ? mt->mass : 0.0;
          dgo[3 * i + j] = sum[3 * i + j] * dt2 + mm;
        }
      /* cg_solver.cuh:323-331 */
      force[0] *= mt->c_bond;
      force[1] *= mt->c_bond;
      force[2] *= mt->c_bond;
      force[3] *= mt->c_bond;
      const double ss0 = (10.0 - 9.0 * clock) * sys->s00[p];
      sys->s_new[p] = ss0 - ORC_ALPHA_PR * smin;
      force[1] -= (clock * prm->gravity * mt->mass);
      force[3] = 0;
      /* contact force (simulator.cuh:762-765), added when non-negligible */
      if (prm->contact_radius > 0.0) {
        const double* c = sys->fc + 4 * (int64_t)p;
        const double len = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2] +
                                c[3] * c[3]);
        if (len > 1e-10) {
          force[0] += c[0];
          force[1] += c[1];
          force[2] += c[2];
          force[3] += c[3];
        }
      }
      double* fo = sys->f + 4 * (int64_t)p;
      fo[0] = force[0];
      fo[1] = force[1];
      fo[2] = force[2];
      fo[3] = force[3];
      /* right-hand side */
      const double* vp = sys->v + 4 * (int64_t)p;
      double* bo = sys->b + 4 * (int64_t)p;
      if (prm->integrator == 1) {
        /* newmark_beta.cuh:82-84 */
        for (int k = 0; k < 4; ++k)
          bo[k] = dt * (mt->mass * vp[k] + 0.5 * dt * force[k]);
      } else {
        /* implicit_euler.cuh:82-84 */
        for (int k = 0; k < 4; ++k) bo[k] = mt->mass * vp[k] + dt * force[k];
      }
      if (top != otop) {
        sys->top[p] = top;
        broken += otop - top;
      }
      live += top + 1;
    }
#pragma omp critical
    {
      if (lmin < gmin || (lmin == gmin && lp >= 0 && lp < gp)) {
        gmin = lmin;
        gp = lp;
        gq = lq;
      }
    }
  }
  if (st) {
    st->broken_bonds = broken;
    st->live_bonds = live;
    st->min_gap = gmin;
    st->gap_p = gp;
    st->gap_q = gq;
  }
}

/* ======================================================================
 * SpMV: core/cg_solver.cuh:343-383 computeAx
 *   y_p = A_pp x_p + sum_b A_{p,b} x_{q_b} (live bonds in list order);
 *   each block product is dot(row, x) = (a0 x0 + a1 x1) + a2 x2 and is
 *   added to the running sum (Apx.x += dot(...)).
 * ====================================================================== */
static inline void blk_mul(const double* A, const double* x, double* o) {
  o[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
  o[1] = A[3] * x[0] + A[4] * x[1] + A[5] * x[2];
  o[2] = A[6] * x[0] + A[7] * x[1] + A[8] * x[2];
}

void orc_spmv(const orc_system* sys, const double* x4, double* y4) {
  const int32_t n = sys->n;
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    double acc[3], t[3];
    blk_mul(sys->diag + 9 * (int64_t)p, x4 + 4 * (int64_t)p, acc);
    const int64_t rs = sys->row_start[p];
    const int32_t top = sys->top[p];
    for (int32_t bnd = 0; bnd <= top; ++bnd) {
      const int32_t q = sys->col[rs + bnd];
      blk_mul(sys->val + 9 * (rs + bnd), x4 + 4 * (int64_t)q, t);
      acc[0] += t[0];
      acc[1] += t[1];
      acc[2] += t[2];
    }
    double* y = y4 + 4 * (int64_t)p;
    y[0] = acc[0];
    y[1] = acc[1];
    y[2] = acc[2];
    y[3] = 0.0;
  }
}

/* y = A x for the solver's first residual: the bond terms summed first,
 * from 0.0 in list order, the diagonal term added last (the order in which
 * the GPU assembly pass can form it while it writes the blocks) */
static void orc_spmv_resid(const orc_system* sys, const double* x4, double* y4) {
  const int32_t n = sys->n;
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    double acc[3] = {0.0, 0.0, 0.0}, t[3], d[3];
    const int64_t rs = sys->row_start[p];
    for (int32_t bnd = 0; bnd <= sys->top[p]; ++bnd) {
      const int32_t q = sys->col[rs + bnd];
      blk_mul(sys->val + 9 * (rs + bnd), x4 + 4 * (int64_t)q, t);
      acc[0] += t[0];
      acc[1] += t[1];
      acc[2] += t[2];
    }
    blk_mul(sys->diag + 9 * (int64_t)p, x4 + 4 * (int64_t)p, d);
    double* y = y4 + 4 * (int64_t)p;
    y[0] = d[0] + acc[0];
    y[1] = d[1] + acc[1];
    y[2] = d[2] + acc[2];
    y[3] = 0.0;
  }
}

/* Inverse of the symmetric 3x3 diagonal block by cofactors; the six
 * distinct entries are computed once and mirrored so the inverse is
 * exactly symmetric. */
void orc_block_jacobi(orc_system* sys) {
  const int32_t n = sys->n;
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    const double* a = sys->diag + 9 * (int64_t)p;
    const double a00 = a[0], a01 = a[1], a02 = a[2];
    const double a11 = a[4], a12 = a[5], a22 = a[8];
    const double c00 = a11 * a22 - a12 * a12;
    const double c01 = a02 * a12 - a01 * a22;
    const double c02 = a01 * a12 - a02 * a11;
    const double c11 = a00 * a22 - a02 * a02;
    const double c12 = a01 * a02 - a00 * a12;
    const double c22 = a00 * a11 - a01 * a01;
    const double det = (a00 * c00 + a01 * c01) + a02 * c02;
    const double id = 1.0 / det;
    double* m = sys->minv + 9 * (int64_t)p;
    m[0] = c00 * id;
    m[1] = c01 * id;
    m[2] = c02 * id;
    m[3] = m[1];
    m[4] = c11 * id;
    m[5] = c12 * id;
    m[6] = m[2];
    m[7] = m[5];
    m[8] = c22 * id;
  }
}

/* ======================================================================
 * CG: CUSP krylov::cg structure (core/simulator.cu:763-770; CUSP is not
 * vendored, the loop below is the textbook PCG of Saad Alg. 9.1 with the
 * CUSP monitor test ||r|| <= rtol ||b|| ahead of every iteration):
 *   y = A x; r = b - y; z = M r; p = z; rz = <r,z>
 *     (this first product sums each row's bond terms first, then the
 *     diagonal term: orc_spmv_resid)
 *   while !(||r|| <= tol or it >= max):
 *     y = A p; alpha = rz / <y,p>; x += alpha p; r -= alpha y;
 *     z = M r; rz' = <r,z>; beta = rz'/rz; p = z + beta p; ++it
 * Dirichlet rows (bc = 1) start at vbc and are projected out of r, so the
 * iteration solves the free rows with the prescribed values fixed.
 * ====================================================================== */
static inline void precond_apply(const orc_system* sys, int precond,
                                 int64_t p, const double* r, double* z) {
  if (precond != 0) { /* 1 block Jacobi, 2 given blocks */
    const double* m = sys->minv + 9 * p;
    z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
    z[1] = m[3] * r[0] + m[4] * r[1] + m[5] * r[2];
    z[2] = m[6] * r[0] + m[7] * r[1] + m[8] * r[2];
  } else {
    z[0] = r[0];
    z[1] = r[1];
    z[2] = r[2];
  }
  z[3] = 0.0;
}

void orc_cg(orc_system* sys, const orc_step_params* prm, orc_step_stats* st) {
  const int32_t n = sys->n;
  const size_t bytes = sizeof(double) * 4 * (size_t)(n > 0 ? n : 1);
  double* x = sys->sol;
  double* y = (double*)malloc(bytes);
  double* r = (double*)malloc(bytes);
  double* z = (double*)malloc(bytes);
  double* p = (double*)malloc(bytes);
  double* sb = (double*)malloc(bytes);
  const double dt = prm->dt;
  if (prm->precond == 1) orc_block_jacobi(sys); /* precond 2: minv as given */

  /* Dirichlet rows start at their prescribed value (implicit Euler solves
   * for the velocity, Newmark-beta for the displacement dt * v). */
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    if (sys->bc[i]) {
      for (int k = 0; k < 4; ++k) {
        const double vb = sys->vbc[4 * (int64_t)i + k];
        x[4 * (int64_t)i + k] = (prm->integrator == 1) ? dt * vb : vb;
      }
      x[4 * (int64_t)i + 3] = 0.0;
    }
  }
  orc_spmv_resid(sys, x, y);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    const int64_t o = 4 * (int64_t)i;
    const int free_row = !sys->bc[i];
    for (int k = 0; k < 3; ++k) {
      r[o + k] = free_row ? (sys->b[o + k] - y[o + k]) : 0.0;
      sb[o + k] = free_row ? sys->b[o + k] : 0.0;
    }
    r[o + 3] = 0.0;
    sb[o + 3] = 0.0;
    precond_apply(sys, prm->precond, i, r + o, z + o);
    for (int k = 0; k < 4; ++k) p[o + k] = z[o + k];
  }
  double rz = orc_dot(n, r, z);
  const double bnorm = sqrt(orc_dot(n, sb, sb));
  const double tol = prm->rtol * bnorm;
  int32_t it = 0;
  int converged = 0;
  double rnorm;
  for (;;) {
    rnorm = sqrt(orc_dot(n, r, r));
    if (rnorm <= tol) {
      converged = 1;
      break;
    }
    if (it >= prm->max_iter) break;
    orc_spmv(sys, p, y);
    const double alpha = rz / orc_dot(n, y, p);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
      const int64_t o = 4 * (int64_t)i;
      for (int k = 0; k < 3; ++k) x[o + k] = alpha * p[o + k] + x[o + k];
      if (!sys->bc[i]) {
        for (int k = 0; k < 3; ++k) r[o + k] = (-alpha) * y[o + k] + r[o + k];
      }
      precond_apply(sys, prm->precond, i, r + o, z + o);
    }
    const double rz_old = rz;
    rz = orc_dot(n, r, z);
    const double beta = rz / rz_old;
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
      const int64_t o = 4 * (int64_t)i;
      for (int k = 0; k < 3; ++k) p[o + k] = z[o + k] + beta * p[o + k];
    }
    ++it;
  }
  if (st) {
    st->cg_iterations = it;
    st->converged = converged;
    st->residual_norm = rnorm;
    st->b_norm = bnorm;
    st->refinements = 0;
  }
  free(y);
  free(r);
  free(z);
  free(p);
  free(sb);
  (void)orc_dot3;
}

/* ======================================================================
 * Mixed-precision solve (precision = 1).  Not in the reference, which
 * solves in FP64 only; this is the textbook iterative refinement (Wilkinson;
 * Higham, "Accuracy and Stability of Numerical Algorithms", ch. 12) with a
 * PCG inner solver:
 *   A32 = the assembled off-diagonal blocks rounded to float, FP64 diagonal
 *   loop:  r = S(b - A x) in FP64 with the FP64 system (orc_spmv_resid:
 *          bond terms summed first, the diagonal term added last)
 *          stop if ||r|| <= rtol ||S b|| (FP64 norms)
 *          inner PCG on float vectors and float arithmetic for A32 d = r,
 *            diagonal block and block-Jacobi inverse rounded to float,
 *            dot products and norms in FP64 (orc_dot's tree over widened
 *            floats), alpha and beta FP64 and rounded to float for the
 *            vector updates, stop when ||r32|| <= max(inner_rtol ||r||,
 *            tol / 2) or after inner_max_iter iterations
 *          x += d (FP64)
 * The FP32 blocks are float((F * -1) * dt2), the same rounding the GPU
 * assembly applies; minv32 = float(minv), minv from orc_block_jacobi.
 * ====================================================================== */
static inline void blk_mul_f(const float* A, const float* x, float* o) {
  o[0] = A[0] * x[0] + A[1] * x[1] + A[2] * x[2];
  o[1] = A[3] * x[0] + A[4] * x[1] + A[5] * x[2];
  o[2] = A[6] * x[0] + A[7] * x[1] + A[8] * x[2];
}

static inline void blk_mul_wide(const float* A, const double* x, double* o) {
  o[0] = (double)A[0] * x[0] + (double)A[1] * x[1] + (double)A[2] * x[2];
  o[1] = (double)A[3] * x[0] + (double)A[4] * x[1] + (double)A[5] * x[2];
  o[2] = (double)A[6] * x[0] + (double)A[7] * x[1] + (double)A[8] * x[2];
}

/* y = A32 x in FP64: FP64 diagonal, float bonds widened, FP64 x; the bond
 * terms are summed first (from 0.0, list order), the diagonal term last */
void orc_spmv_wide(const orc_system* sys, const double* x4, double* y4) {
  const int32_t n = sys->n;
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    double acc[3] = {0.0, 0.0, 0.0}, t[3], d[3];
    const int64_t rs = sys->row_start[p];
    for (int32_t bnd = 0; bnd <= sys->top[p]; ++bnd) {
      const int32_t q = sys->col[rs + bnd];
      blk_mul_wide(sys->val32 + 9 * (rs + bnd), x4 + 4 * (int64_t)q, t);
      acc[0] += t[0];
      acc[1] += t[1];
      acc[2] += t[2];
    }
    blk_mul(sys->diag + 9 * (int64_t)p, x4 + 4 * (int64_t)p, d);
    double* y = y4 + 4 * (int64_t)p;
    y[0] = d[0] + acc[0];
    y[1] = d[1] + acc[1];
    y[2] = d[2] + acc[2];
    y[3] = 0.0;
  }
}

/* y = A32 x in float arithmetic, same order as orc_spmv */
static void orc_spmv_f32(const orc_system* sys, const float* x4, float* y4) {
  const int32_t n = sys->n;
#pragma omp parallel for schedule(static)
  for (int32_t p = 0; p < n; ++p) {
    float acc[3], t[3];
    blk_mul_f(sys->diag32 + 9 * (int64_t)p, x4 + 4 * (int64_t)p, acc);
    const int64_t rs = sys->row_start[p];
    for (int32_t bnd = 0; bnd <= sys->top[p]; ++bnd) {
      const int32_t q = sys->col[rs + bnd];
      blk_mul_f(sys->val32 + 9 * (rs + bnd), x4 + 4 * (int64_t)q, t);
      acc[0] += t[0];
      acc[1] += t[1];
      acc[2] += t[2];
    }
    float* y = y4 + 4 * (int64_t)p;
    y[0] = acc[0];
    y[1] = acc[1];
    y[2] = acc[2];
    y[3] = 0.0f;
  }
}

static inline void precond_apply_f(const orc_system* sys, int precond, int64_t p,
                                   const float* r, float* z) {
  if (precond != 0) {
    const float* m = sys->minv32 + 9 * p;
    z[0] = m[0] * r[0] + m[1] * r[1] + m[2] * r[2];
    z[1] = m[3] * r[0] + m[4] * r[1] + m[5] * r[2];
    z[2] = m[6] * r[0] + m[7] * r[1] + m[8] * r[2];
  } else {
    z[0] = r[0];
    z[1] = r[1];
    z[2] = r[2];
  }
  z[3] = 0.0f;
}

void orc_mixed_matrix(orc_system* sys) {
  orc_block_jacobi(sys);
  orc_round_blocks(sys);
}

void orc_round_blocks(orc_system* sys) {
  const int32_t n = sys->n;
  const int64_t nslot = sys->row_start[n];
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < 9 * nslot; ++k) sys->val32[k] = (float)sys->val[k];
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < 9 * (int64_t)n; ++k) {
    sys->diag32[k] = (float)sys->diag[k];
    sys->minv32[k] = (float)sys->minv[k];
  }
}

void orc_cg_mixed(orc_system* sys, const orc_step_params* prm, orc_step_stats* st) {
  const int32_t n = sys->n;
  const size_t nn = (size_t)(n > 0 ? n : 1);
  double* x = sys->sol;
  double* y = (double*)malloc(sizeof(double) * 4 * nn);
  float* r32 = (float*)malloc(sizeof(float) * 4 * nn);
  float* z32 = (float*)malloc(sizeof(float) * 4 * nn);
  float* p32 = (float*)malloc(sizeof(float) * 4 * nn);
  float* y32 = (float*)malloc(sizeof(float) * 4 * nn);
  float* d32 = (float*)malloc(sizeof(float) * 4 * nn);
  double* t_rr = (double*)malloc(sizeof(double) * nn);
  double* t_rz = (double*)malloc(sizeof(double) * nn);
  double* t_r32 = (double*)malloc(sizeof(double) * nn);
  double* t_bb = (double*)malloc(sizeof(double) * nn);
  const double dt = prm->dt;
  if (prm->precond == 2)
    orc_round_blocks(sys); /* minv as given */
  else
    orc_mixed_matrix(sys);
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    if (sys->bc[i]) {
      for (int k = 0; k < 3; ++k) {
        const double vb = sys->vbc[4 * (int64_t)i + k];
        x[4 * (int64_t)i + k] = (prm->integrator == 1) ? dt * vb : vb;
      }
      x[4 * (int64_t)i + 3] = 0.0;
    }
  }
  int32_t it = 0, ref = 0, converged = 0, first = 1;
  double bnorm = 0.0, tol = 0.0, rnorm = 0.0;
  for (;;) {
    /* FP64 residual of the FP64 system, then the inner solve's start:
     * d = 0, z = M r, p = z */
    orc_spmv_resid(sys, x, y);
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
      const int64_t o = 4 * (int64_t)i;
      const int fr = !sys->bc[i];
      double r[3], sb[3];
      for (int k = 0; k < 3; ++k) {
        r[k] = fr ? (sys->b[o + k] - y[o + k]) : 0.0;
        sb[k] = fr ? sys->b[o + k] : 0.0;
        r32[o + k] = (float)r[k];
        d32[o + k] = 0.0f;
      }
      r32[o + 3] = 0.0f;
      d32[o + 3] = 0.0f;
      precond_apply_f(sys, prm->precond, i, r32 + o, z32 + o);
      for (int k = 0; k < 4; ++k) p32[o + k] = z32[o + k];
      t_rr[i] = term3(r, r);
      t_bb[i] = term3(sb, sb);
      t_rz[i] = term3f(r32 + o, z32 + o);
      t_r32[i] = term3f(r32 + o, r32 + o);
    }
    const double rr = orc_tree(n, t_rr);
    double rz = orc_tree(n, t_rz);
    double rr32 = orc_tree(n, t_r32);
    const double bb = orc_tree(n, t_bb);
    if (first) {
      bnorm = sqrt(bb);
      tol = prm->rtol * bnorm;
      first = 0;
    }
    rnorm = sqrt(rr);
    if (rnorm <= tol) {
      converged = 1;
      break;
    }
    if (ref >= prm->max_refine || it >= prm->max_iter) break;
    ++ref;
    const double a_tol = prm->inner_rtol * rnorm, b_tol = 0.5 * tol;
    const double inner_tol = a_tol > b_tol ? a_tol : b_tol;
    int32_t k_in = 0;
    for (;;) {
      if (sqrt(rr32) <= inner_tol || k_in >= prm->inner_max_iter || it >= prm->max_iter) break;
      orc_spmv_f32(sys, p32, y32);
      const double alpha = rz / orc_dot_f32(n, y32, p32);
      const float af = (float)alpha;
#pragma omp parallel for schedule(static)
      for (int32_t i = 0; i < n; ++i) {
        const int64_t o = 4 * (int64_t)i;
        for (int k = 0; k < 3; ++k) d32[o + k] = af * p32[o + k] + d32[o + k];
        if (!sys->bc[i]) {
          for (int k = 0; k < 3; ++k) r32[o + k] = (-af) * y32[o + k] + r32[o + k];
        }
        precond_apply_f(sys, prm->precond, i, r32 + o, z32 + o);
        t_rz[i] = term3f(r32 + o, z32 + o);
        t_r32[i] = term3f(r32 + o, r32 + o);
      }
      const double rz_new = orc_tree(n, t_rz);
      rr32 = orc_tree(n, t_r32);
      const double beta = rz_new / rz;
      rz = rz_new;
      const float bf = (float)beta;
      ++k_in;
      ++it;
      if (sqrt(rr32) <= inner_tol || k_in >= prm->inner_max_iter || it >= prm->max_iter) break;
#pragma omp parallel for schedule(static)
      for (int32_t i = 0; i < n; ++i) {
        const int64_t o = 4 * (int64_t)i;
        for (int k = 0; k < 3; ++k) p32[o + k] = z32[o + k] + bf * p32[o + k];
      }
    }
    /* x += d in FP64 */
#pragma omp parallel for schedule(static)
    for (int32_t i = 0; i < n; ++i) {
      const int64_t o = 4 * (int64_t)i;
      for (int k = 0; k < 3; ++k) x[o + k] = x[o + k] + (double)d32[o + k];
    }
  }
  if (st) {
    st->cg_iterations = it;
    st->converged = converged;
    st->residual_norm = rnorm;
    st->b_norm = bnorm;
    st->refinements = ref;
  }
  free(y);
  free(r32);
  free(z32);
  free(p32);
  free(y32);
  free(d32);
  free(t_rr);
  free(t_rz);
  free(t_r32);
  free(t_bb);
}

/* ======================================================================
 * Position update with wall clamp: core/implicit_euler.cuh:90-176
 * (integrator 0: v <- sol, x += v dt) and core/newmark_beta.cuh:89-199
 * (integrator 1: v <- 2 dx / dt - v, x += dx).
 * ====================================================================== */
static inline void clamp_axis(double* pos, double* vel, double lo, double hi,
                              double rest) {
  if (*pos > hi) {
    *vel *= -rest;
    *pos = hi - rest * (*pos - hi);
  }
  if (*pos < lo) {
    *vel *= -rest;
    *pos = lo + rest * (lo - *pos);
  }
}

void orc_update(orc_system* sys, const orc_step_params* prm) {
  const int32_t n = sys->n;
  const double dt = prm->dt;
#pragma omp parallel for schedule(static)
  for (int32_t i = 0; i < n; ++i) {
    if (sys->act[i] == ORC_INVISIBLE) continue;
    const int64_t o = 4 * (int64_t)i;
    double* xp = sys->x + o;
    double* vp = sys->v + o;
    const double* s = sys->sol + o;
    double nv[4], np[4];
    if (prm->integrator == 1) {
      for (int k = 0; k < 4; ++k) nv[k] = 2.0 / dt * s[k] - vp[k];
      for (int k = 0; k < 4; ++k) np[k] = xp[k] + s[k];
    } else {
      for (int k = 0; k < 4; ++k) nv[k] = s[k];
      for (int k = 0; k < 4; ++k) np[k] = xp[k] + nv[k] * dt;
    }
    if (prm->walls) {
      const double r = prm->wall_radius;
      for (int k = 0; k < 3; ++k)
        clamp_axis(&np[k], &nv[k], prm->bmin[k] + r, prm->bmax[k] - r,
                   prm->restitution);
    }
    for (int k = 0; k < 4; ++k) {
      xp[k] = np[k];
      vp[k] = nv[k];
    }
  }
}

/* ======================================================================
 * Contact: core/simulator.cuh:679-766 collidePDParticles, restated on the
 * uniform grid shared with the HIP path:
 *   cell edge h_c = contact_cell, origin = componentwise min of x,
 *   cell (ix,iy,iz) = floor((x - origin) / h_c) per axis,
 *   dims = floor((max - origin) / h_c) + 1, key = (iz*ny + iy)*nx + ix.
 * Each particle visits the 27 neighbour cells in z, y, x order (-1..1
 * each), members of a cell in ascending particle index, and accumulates
 *   force += (Kc (r - Rc)^2 / V) * ((x_p - x_q) / r)   for 0 < r < Rc
 * over particles q of another body.
 * ====================================================================== */
void orc_contact(orc_system* sys, const orc_step_params* prm) {
  const int32_t n = sys->n;
  memset(sys->fc, 0, sizeof(double) * 4 * (size_t)n);
  if (prm->contact_radius <= 0.0 || n == 0) return;
  const double hc = prm->contact_cell;
  double mn[3] = {INFINITY, INFINITY, INFINITY};
  double mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      const double c = sys->x[4 * (int64_t)i + k];
      if (c < mn[k]) mn[k] = c;
      if (c > mx[k]) mx[k] = c;
    }
  int64_t dims[3];
  for (int k = 0; k < 3; ++k) dims[k] = (int64_t)floor((mx[k] - mn[k]) / hc) + 1;
  const int64_t ncell = dims[0] * dims[1] * dims[2];
  int64_t* key = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
  int64_t* start = (int64_t*)calloc((size_t)ncell + 1, sizeof(int64_t));
  int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
  for (int32_t i = 0; i < n; ++i) {
    int64_t c[3];
    for (int k = 0; k < 3; ++k)
      c[k] = (int64_t)floor((sys->x[4 * (int64_t)i + k] - mn[k]) / hc);
    key[i] = (c[2] * dims[1] + c[1]) * dims[0] + c[0];
    start[key[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  {
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)ncell);
    memcpy(fill, start, sizeof(int64_t) * (size_t)ncell);
    for (int32_t i = 0; i < n; ++i) sorted[fill[key[i]]++] = i; /* stable */
    free(fill);
  }
  const double Rc = prm->contact_radius;
  const double kc = prm->contact_stiffness;
  const double vol = prm->contact_volume;
#pragma omp parallel for schedule(dynamic, 256)
  for (int32_t p = 0; p < n; ++p) {
    if (sys->act[p] == ORC_INVISIBLE) continue;
    const double* pp = sys->x + 4 * (int64_t)p;
    int64_t c[3];
    for (int k = 0; k < 3; ++k) c[k] = (int64_t)floor((pp[k] - mn[k]) / hc);
    double force[4] = {0, 0, 0, 0};
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int64_t cx = c[0] + dx, cy = c[1] + dy, cz = c[2] + dz;
          if (cx < 0 || cy < 0 || cz < 0 || cx >= dims[0] || cy >= dims[1] ||
              cz >= dims[2])
            continue;
          const int64_t cell = (cz * dims[1] + cy) * dims[0] + cx;
          for (int64_t m = start[cell]; m < start[cell + 1]; ++m) {
            const int32_t q = sorted[m];
            if (q == p) continue;
            if (prm->contact_mode == 0 && sys->body[q] == sys->body[p]) continue;
            if (sys->act[q] == ORC_INVISIBLE) continue;
            const double* xq = sys->x + 4 * (int64_t)q;
            const double pq[4] = {pp[0] - xq[0], pp[1] - xq[1], pp[2] - xq[2],
                                  pp[3] - xq[3]};
            const double r =
                sqrt(pq[0] * pq[0] + pq[1] * pq[1] + pq[2] * pq[2]);
            if (r >= Rc) continue;
            if (r == 0.0) continue;
            const double howfar = (r - Rc);
            const double mag = kc * (howfar * howfar);
            const double coef = mag / vol;
            for (int k = 0; k < 4; ++k) force[k] += coef * (pq[k] / r);
          }
        }
    double* o = sys->fc + 4 * (int64_t)p;
    for (int k = 0; k < 4; ++k) o[k] = force[k];
  }
  free(key);
  free(start);
  free(sorted);
}

/* core/simulator.cu:297-467 (implicit Euler) / 470-652 (Newmark-beta)
 * without the adaptive-rollback branch, which the host Simulator owns. */
int32_t orc_step(orc_system* sys, const orc_step_params* prm,
                 orc_step_stats* st) {
  orc_contact(sys, prm);
  orc_assemble(sys, prm, st);
  /* simulator.cuh:2612 s <- s_new */
  memcpy(sys->s, sys->s_new, sizeof(double) * (size_t)sys->n);
  if (prm->precision == 1)
    orc_cg_mixed(sys, prm, st);
  else
    orc_cg(sys, prm, st);
  orc_update(sys, prm);
  return (st && !st->converged) ? 1 : 0;
}
