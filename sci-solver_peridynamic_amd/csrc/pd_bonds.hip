// pd_bonds.hip -- uniform-grid cell hashing, LDS radix sort, cell
// start/end and the horizon bond-list build.
//
// Replaces the reference's host kd-tree (core/kdtree.h, core/simulator.cu:
// 825-943) and reuses the grid / sort / cell-range scheme of its SPH path
// (core/simulator.cuh:61-88, 862-887, 989-1062; core/simulator.cu:1019-1051)
// for the GPU.  The neighbour predicate is the kd-tree's distance
// sqrt(((0 + dx^2) + dy^2) + dz^2) < horizon with dx = x_p - x_q, so the
// lists equal the reference's bit for bit once sorted.
#include <math.h>

#include "pd_internal.h"
#include "pd_solver.h"  // PD_MAX_BODIES

namespace pd {

// ---------------------------------------------------------------- bbox
// Bounding box of the particles of one body (body_sel >= 0) or of all.
__global__ __launch_bounds__(PD_BLOCK) void k_bbox_partial(const double4* x, const int32_t* body,
                                                          int n, int body_sel, double* part) {
  __shared__ double s[6][PD_BLOCK];
  double lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY;
  double hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
  for (int i = blockIdx.x * PD_BLOCK + threadIdx.x; i < n; i += gridDim.x * PD_BLOCK) {
    if (body_sel >= 0 && body[i] != body_sel) continue;
    const double4 p = x[i];
    lo0 = fmin(lo0, p.x);
    lo1 = fmin(lo1, p.y);
    lo2 = fmin(lo2, p.z);
    hi0 = fmax(hi0, p.x);
    hi1 = fmax(hi1, p.y);
    hi2 = fmax(hi2, p.z);
  }
  s[0][threadIdx.x] = lo0;
  s[1][threadIdx.x] = lo1;
  s[2][threadIdx.x] = lo2;
  s[3][threadIdx.x] = hi0;
  s[4][threadIdx.x] = hi1;
  s[5][threadIdx.x] = hi2;
  __syncthreads();
  for (int w = PD_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
      for (int k = 3; k < 6; ++k) s[k][threadIdx.x] = fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(PD_BLOCK) void k_bbox_final(const double* part, int g, double* out6) {
  __shared__ double s[6][PD_BLOCK];
  for (int k = 0; k < 6; ++k) {
    double v = k < 3 ? INFINITY : -INFINITY;
    for (int b = threadIdx.x; b < g; b += PD_BLOCK)
      v = k < 3 ? fmin(v, part[b * 6 + k]) : fmax(v, part[b * 6 + k]);
    s[k][threadIdx.x] = v;
  }
  __syncthreads();
  for (int w = PD_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
      for (int k = 3; k < 6; ++k) s[k][threadIdx.x] = fmax(s[k][threadIdx.x], s[k][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 6) out6[threadIdx.x] = s[threadIdx.x][0];
}

static const int kBboxBlocks = 256;
// all-body boxes: four workgroups per CU keep the position stream in flight
static const int kBodyBboxBlocks = 1024;

// std::min / std::max as the host forms them (the first operand on ties,
// so -0 and +0 keep their order)
__device__ __forceinline__ double first_min(double a, double b) { return b < a ? b : a; }
__device__ __forceinline__ double first_max(double a, double b) { return a < b ? b : a; }

// The contact plan (one thread): the zones are the body boxes widened by
// w, and the sub-grid box is the bound of every zone's intersection with
// another body's box, formed in the order of the host loops it replaces.
__device__ void contact_plan(const double* bb, int nb, double w, ContactPlan& P) {
  for (int k = 0; k < 3; ++k) {
    P.mn[k] = P.ulo[k] = P.glo[k] = INFINITY;
    P.mx[k] = P.uhi[k] = P.ghi[k] = -INFINITY;
  }
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) {
      P.mn[k] = first_min(P.mn[k], bb[6 * b + k]);
      P.mx[k] = first_max(P.mx[k], bb[6 * b + 3 + k]);
    }
  P.z.nbody = nb;
  for (int b = 0; b < PD_MAX_ZONE_BODIES; ++b)
    for (int k = 0; k < 3; ++k) {
      P.z.lo[b][k] = b < nb ? bb[6 * b + k] - w : INFINITY;  // empty bodies: +inf, no zone
      P.z.hi[b][k] = b < nb ? bb[6 * b + 3 + k] + w : -INFINITY;
    }
  for (int b = 0; b < nb; ++b)
    for (int k = 0; k < 3; ++k) {
      P.ulo[k] = first_min(P.ulo[k], P.z.lo[b][k]);
      P.uhi[k] = first_max(P.uhi[k], P.z.hi[b][k]);
    }
  for (int b = 0; b < nb; ++b)
    for (int o = 0; o < nb; ++o) {
      if (o == b) continue;
      double lo[3], hi[3];
      bool meet = true;
      for (int k = 0; k < 3; ++k) {
        lo[k] = first_max(P.z.lo[b][k], bb[6 * o + k]);
        hi[k] = first_min(P.z.hi[b][k], bb[6 * o + 3 + k]);
        meet = meet && lo[k] <= hi[k];
      }
      if (!meet) continue;
      for (int k = 0; k < 3; ++k) {
        P.glo[k] = first_min(P.glo[k], lo[k]);
        P.ghi[k] = first_max(P.ghi[k], hi[k]);
      }
    }
}

// Bounding boxes of all bodies in one pass: per block (lo xyz, hi xyz) for
// each body b < PD_MAX_BODIES at part[(block * PD_MAX_BODIES + b) * 6 + k].
// min / max are exact, so the result does not depend on the order.
__global__ __launch_bounds__(PD_BLOCK) void k_bbox_bodies_partial(const double4* x,
                                                                 const int32_t* body, int n,
                                                                 int nbody, double* part) {
  double v[PD_MAX_BODIES][6];
#pragma unroll
  for (int b = 0; b < PD_MAX_BODIES; ++b)
#pragma unroll
    for (int k = 0; k < 6; ++k) v[b][k] = k < 3 ? INFINITY : -INFINITY;
  for (int i = blockIdx.x * PD_BLOCK + threadIdx.x; i < n; i += gridDim.x * PD_BLOCK) {
    const int bi = body[i];
    const double4 p = x[i];
#pragma unroll
    for (int b = 0; b < PD_MAX_BODIES; ++b)
      if (b == bi) {
        v[b][0] = fmin(v[b][0], p.x);
        v[b][1] = fmin(v[b][1], p.y);
        v[b][2] = fmin(v[b][2], p.z);
        v[b][3] = fmax(v[b][3], p.x);
        v[b][4] = fmax(v[b][4], p.y);
        v[b][5] = fmax(v[b][5], p.z);
      }
  }
  __shared__ double s[PD_BLOCK / 64][PD_MAX_BODIES * 6];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int b = 0; b < PD_MAX_BODIES; ++b) {
    if (b >= nbody) break;  // uniform: only the bodies that exist are folded
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      double t = v[b][k];
#pragma unroll
      for (int m = 32; m >= 1; m >>= 1) {
        const double o = __shfl_xor(t, m, 64);
        t = k < 3 ? fmin(t, o) : fmax(t, o);
      }
      if (lane == 0) s[wave][b * 6 + k] = t;
    }
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t < nbody * 6) {
    const bool lo = (t % 6) < 3;
    double r = s[0][t];
    for (int w = 1; w < PD_BLOCK / 64; ++w) r = lo ? fmin(r, s[w][t]) : fmax(r, s[w][t]);
    part[blockIdx.x * PD_MAX_BODIES * 6 + t] = r;
  }
}

// each thread folds the block partials t, t + 256, ... (g <= 4 PD_BLOCK),
// then every value of the bodies that exist is folded by wave butterflies
// and across the four waves; the other bodies' entries are left unwritten.
// plan non-null: thread 0 then forms the contact plan from the boxes
// (contact_plan, zones widened by w).
__global__ __launch_bounds__(PD_BLOCK) void k_bbox_bodies_final(const double* part, int g,
                                                               int nbody, double* out,
                                                               ContactPlan* plan, double w) {
  __shared__ double s[PD_BLOCK / 64][PD_MAX_BODIES * 6];
  __shared__ double box[PD_MAX_BODIES * 6];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  for (int k = 0; k < nbody * 6; ++k) {
    const bool lo = (k % 6) < 3;
    double r = lo ? INFINITY : -INFINITY;
    for (int i = t; i < g; i += PD_BLOCK) {
      const double v = part[i * PD_MAX_BODIES * 6 + k];
      r = lo ? fmin(r, v) : fmax(r, v);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      const double o = __shfl_xor(r, m, 64);
      r = lo ? fmin(r, o) : fmax(r, o);
    }
    if (lane == 0) s[wave][k] = r;
  }
  __syncthreads();
  if (t < nbody * 6) {
    const bool lo = (t % 6) < 3;
    double r = s[0][t];
    for (int w = 1; w < PD_BLOCK / 64; ++w) r = lo ? fmin(r, s[w][t]) : fmax(r, s[w][t]);
    out[t] = r;
    box[t] = r;
  }
  if (!plan) return;  // uniform
  __shared__ ContactPlan sp;
  __syncthreads();
  if (t == 0) contact_plan(box, nbody, w, sp);
  __syncthreads();
  static_assert(sizeof(ContactPlan) % 8 == 0, "the plan is copied in 8-byte words");
  const uint64_t* src = (const uint64_t*)&sp;
  uint64_t* dst = (uint64_t*)plan;
  for (int i = t; i < (int)(sizeof(ContactPlan) / 8); i += PD_BLOCK) dst[i] = src[i];
}

void launch_bbox_bodies(const double4* x, const int32_t* body, int32_t n, int32_t nbody,
                        double* part, double* out, hipStream_t s, ContactPlan* plan, double w) {
  hipLaunchKernelGGL(k_bbox_bodies_partial, dim3(kBodyBboxBlocks), dim3(PD_BLOCK), 0, s, x,
                     body, n, nbody, part);
  hipLaunchKernelGGL(k_bbox_bodies_final, dim3(1), dim3(PD_BLOCK), 0, s, part, kBodyBboxBlocks,
                     nbody, out, plan, w);
}

void launch_bbox(const double4* x, const int32_t* body, int32_t n, int32_t body_sel,
                 double* part, double* out6, hipStream_t s) {
  hipLaunchKernelGGL(k_bbox_partial, dim3(kBboxBlocks), dim3(PD_BLOCK), 0, s, x, body, n,
                     body_sel, part);
  hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(PD_BLOCK), 0, s, part, kBboxBlocks, out6);
}

// ------------------------------------------------------------ compaction
// Stable selection of the particles inside any zone box of another body:
// per-tile counts, an exclusive scan of the counts, then each tile writes
// its selected indices in ascending order.
#define CP_ITEMS 4
#define CP_TILE (PD_BLOCK * CP_ITEMS)

__device__ __forceinline__ bool in_zone(const double4 p, int b, const ZoneBoxes* __restrict__ z) {
  for (int k = 0; k < z->nbody; ++k) {
    if (k == b) continue;
    if (p.x >= z->lo[k][0] && p.x <= z->hi[k][0] && p.y >= z->lo[k][1] &&
        p.y <= z->hi[k][1] && p.z >= z->lo[k][2] && p.z <= z->hi[k][2])
      return true;
  }
  return false;
}

__global__ __launch_bounds__(PD_BLOCK) void k_zone_count(const double4* x, const int32_t* body,
                                                        int n, const ZoneBoxes* z,
                                                        uint32_t* tile_cnt, double4* fc) {
  __shared__ uint32_t c;
  if (threadIdx.x == 0) c = 0;
  __syncthreads();
  uint32_t mine = 0;
  for (int j = 0; j < CP_ITEMS; ++j) {
    const int i = blockIdx.x * CP_TILE + j * PD_BLOCK + threadIdx.x;
    if (i < n) {
      if (fc) fc[i] = make_double4(0.0, 0.0, 0.0, 0.0);
      if (in_zone(x[i], body[i], z)) ++mine;
    }
  }
  if (mine) atomicAdd(&c, mine);
  __syncthreads();
  if (threadIdx.x == 0) tile_cnt[blockIdx.x] = c;
}

// exclusive scan of g tile counts (one workgroup); total at tile_cnt[g]
__global__ __launch_bounds__(1024) void k_scan_small(uint32_t* v, int g) {
  __shared__ uint32_t sums[1024];
  const int t = threadIdx.x;
  const int chunk = (g + 1023) / 1024;
  const int lo = t * chunk;
  const int hi = min(lo + chunk, g);
  uint32_t acc = 0;
  for (int i = lo; i < hi; ++i) acc += v[i];
  sums[t] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t w = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += w;
    __syncthreads();
  }
  uint32_t run = t > 0 ? sums[t - 1] : 0u;
  for (int i = lo; i < hi; ++i) {
    const uint32_t c = v[i];
    v[i] = run;
    run += c;
  }
  if (t == 1023) v[g] = sums[1023];
}

__global__ __launch_bounds__(PD_BLOCK) void k_zone_scatter(const double4* x, const int32_t* body,
                                                          int n, const ZoneBoxes* z,
                                                          const uint32_t* tile_off,
                                                          int32_t* out) {
  __shared__ uint32_t wcnt[4];
  __shared__ uint32_t run;
  const int w = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  if (threadIdx.x == 0) run = tile_off[blockIdx.x];
  for (int j = 0; j < CP_ITEMS; ++j) {
    const int i = blockIdx.x * CP_TILE + j * PD_BLOCK + threadIdx.x;
    const bool sel = i < n && in_zone(x[i], body[i], z);
    const unsigned long long m = __ballot(sel);
    __syncthreads();
    if (lane == 0) wcnt[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t pos = run + (uint32_t)__popcll(m & lt);
    for (int ww = 0; ww < w; ++ww) pos += wcnt[ww];
    if (sel) out[pos] = i;
    __syncthreads();
    if (threadIdx.x == 0) run += wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
  }
}

void launch_zone_select(const double4* x, const int32_t* body, int32_t n, const ZoneBoxes* z,
                        uint32_t* tile_cnt, int32_t* out, hipStream_t s, double4* fc) {
  const int g = (n + CP_TILE - 1) / CP_TILE;
  if (g == 0) return;
  hipLaunchKernelGGL(k_zone_count, dim3(g), dim3(PD_BLOCK), 0, s, x, body, n, z, tile_cnt, fc);
  hipLaunchKernelGGL(k_scan_small, dim3(1), dim3(1024), 0, s, tile_cnt, g);
  hipLaunchKernelGGL(k_zone_scatter, dim3(g), dim3(PD_BLOCK), 0, s, x, body, n, z, tile_cnt, out);
}

size_t zone_tile_words(int32_t n) { return (size_t)(n + CP_TILE - 1) / CP_TILE + 1; }

// ---------------------------------------------------------------- keys
__device__ __forceinline__ void cell_of(const double4 p, const double* mn, double cell,
                                        int& cx, int& cy, int& cz) {
  cx = (int)floor((p.x - mn[0]) / cell);
  cy = (int)floor((p.y - mn[1]) / cell);
  cz = (int)floor((p.z - mn[2]) / cell);
}

struct KeyArgs {
  const double4* x;
  const int32_t* subset;  // particle of entry i (null: entry i is particle i)
  int n;
  double mn[3];
  double cell;
  int c0[3];
  int dims[3];
  uint32_t* keys;
  int32_t* idx;
};

__global__ __launch_bounds__(PD_BLOCK) void k_cell_keys(KeyArgs a) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i >= a.n) return;
  const int p = a.subset ? a.subset[i] : i;
  int cx, cy, cz;
  cell_of(a.x[p], a.mn, a.cell, cx, cy, cz);
  cx -= a.c0[0];
  cy -= a.c0[1];
  cz -= a.c0[2];
  a.keys[i] = (uint32_t)(((int64_t)cz * a.dims[1] + cy) * a.dims[0] + cx);
  a.idx[i] = p;
}

void launch_cell_keys(const double4* x, const int32_t* subset, int32_t n, const double* mn,
                      double cell, const int32_t* c0, const int32_t* dims, uint32_t* keys,
                      int32_t* idx, hipStream_t s) {
  KeyArgs a;
  a.x = x;
  a.subset = subset;
  a.n = n;
  for (int k = 0; k < 3; ++k) {
    a.mn[k] = mn[k];
    a.dims[k] = dims[k];
    a.c0[k] = c0 ? c0[k] : 0;
  }
  a.cell = cell;
  a.keys = keys;
  a.idx = idx;
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_cell_keys, dim3(nb), dim3(PD_BLOCK), 0, s, a);
}

// ------------------------------------------------------------- radix sort
// Stable LSD radix sort, 8-bit digits.  A workgroup owns a tile of
// 256 x 16 keys: (1) LDS histogram of the tile, (2) one-workgroup exclusive
// scan of the digit-major histogram table, (3) stable scatter: the tile is
// walked in 16 rounds of 256 keys; inside a wave the lanes holding the same
// digit are found with eight ballots, counted per wave in LDS, and each key
// lands at base[digit] + keys of that digit earlier in the tile.
#define RS_ITEMS 16
#define RS_TILE (PD_BLOCK * RS_ITEMS)

size_t radix_hist_words(int32_t n) {
  const size_t nb = (size_t)((n + RS_TILE - 1) / RS_TILE);
  return 256 * (nb > 0 ? nb : 1);
}

__global__ __launch_bounds__(PD_BLOCK) void k_radix_hist(const uint32_t* keys, int n, int shift,
                                                        uint32_t* hist, int nb) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int base = blockIdx.x * RS_TILE;
  for (int j = 0; j < RS_ITEMS; ++j) {
    const int i = base + j * PD_BLOCK + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of hist[0..total) by one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_radix_scan(uint32_t* hist, int total) {
  __shared__ uint32_t sums[1024];
  const int t = threadIdx.x;
  const int chunk = (total + 1023) / 1024;
  const int lo = t * chunk;
  const int hi = min(lo + chunk, total);
  uint32_t acc = 0;
  for (int i = lo; i < hi; ++i) acc += hist[i];
  sums[t] = acc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const uint32_t v = t >= off ? sums[t - off] : 0u;
    __syncthreads();
    sums[t] += v;
    __syncthreads();
  }
  uint32_t run = t > 0 ? sums[t - 1] : 0u;
  for (int i = lo; i < hi; ++i) {
    const uint32_t c = hist[i];
    hist[i] = run;
    run += c;
  }
}

__global__ __launch_bounds__(PD_BLOCK) void k_radix_scatter(const uint32_t* kin,
                                                           const int32_t* vin,
                                                           uint32_t* kout, int32_t* vout,
                                                           int n, int shift,
                                                           const uint32_t* hist, int nb) {
  __shared__ uint32_t run[256];
  __shared__ uint32_t base[256];
  __shared__ uint32_t wcnt[4][256];
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  run[t] = 0;
  base[t] = hist[(int64_t)t * nb + blockIdx.x];
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int r = 0; r < RS_ITEMS; ++r) {
    const int i = blockIdx.x * RS_TILE + r * PD_BLOCK + t;
    const bool valid = i < n;
    const uint32_t k = valid ? kin[i] : 0u;
    const int32_t v = valid ? vin[i] : 0;
    const uint32_t d = (k >> shift) & 255u;
    wcnt[0][t] = 0;
    wcnt[1][t] = 0;
    wcnt[2][t] = 0;
    wcnt[3][t] = 0;
    __syncthreads();
    unsigned long long peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const unsigned long long m = __ballot(bit);
      peers &= bit ? m : ~m;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = base[d] + run[d] + rank;
      for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    run[t] += wcnt[0][t] + wcnt[1][t] + wcnt[2][t] + wcnt[3][t];
    __syncthreads();
  }
}

int radix_sort(uint32_t* keys_a, int32_t* idx_a, uint32_t* keys_b, int32_t* idx_b,
               int32_t n, int32_t bits, uint32_t* hist, hipStream_t s) {
  const int nb = (n + RS_TILE - 1) / RS_TILE;
  if (n <= 0 || nb == 0) return 0;
  const int passes = (bits + 7) / 8;
  uint32_t* kin = keys_a;
  int32_t* vin = idx_a;
  uint32_t* kout = keys_b;
  int32_t* vout = idx_b;
  int where = 0;
  for (int ps = 0; ps < passes; ++ps) {
    const int shift = 8 * ps;
    hipLaunchKernelGGL(k_radix_hist, dim3(nb), dim3(PD_BLOCK), 0, s, kin, n, shift, hist, nb);
    hipLaunchKernelGGL(k_radix_scan, dim3(1), dim3(1024), 0, s, hist, 256 * nb);
    hipLaunchKernelGGL(k_radix_scatter, dim3(nb), dim3(PD_BLOCK), 0, s, kin, vin, kout, vout, n,
                       shift, hist, nb);
    uint32_t* tk = kin;
    kin = kout;
    kout = tk;
    int32_t* tv = vin;
    vin = vout;
    vout = tv;
    where ^= 1;
  }
  return where;
}

// ------------------------------------------------------------ cell ranges
__global__ __launch_bounds__(PD_BLOCK) void k_fill2(int32_t* a, int32_t* b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i < n) {
    a[i] = 0;
    b[i] = 0;
  }
}

__global__ __launch_bounds__(PD_BLOCK) void k_cell_ranges(const uint32_t* keys, int n,
                                                         int32_t* start, int32_t* end) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i >= n) return;
  const uint32_t k = keys[i];
  if (i == 0 || keys[i - 1] != k) start[k] = i;
  if (i == n - 1 || keys[i + 1] != k) end[k] = i + 1;
}

void launch_cell_ranges(const uint32_t* keys, int32_t n, int32_t* start, int32_t* end,
                        int64_t ncell, hipStream_t s) {
  const int64_t nbc = (ncell + PD_BLOCK - 1) / PD_BLOCK;
  if (nbc) hipLaunchKernelGGL(k_fill2, dim3((unsigned)nbc), dim3(PD_BLOCK), 0, s, start, end, ncell);
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_cell_ranges, dim3(nb), dim3(PD_BLOCK), 0, s, keys, n, start, end);
}

// ------------------------------------------------------- small sub-grids
// The per-step contact grid holds a few thousand candidates over a few
// thousand cells.  One 1024-thread workgroup then does the whole build that
// k_cell_keys, the radix passes and the cell-range kernels do in eleven or
// more launches: cell start / end cleared, keys, the same stable LSD radix
// sort (8-bit digits; each chunk of 1024 keys ranked by digit with eight
// ballots per wave and per-wave digit counts in LDS), then the ranges.
// Results equal the multi-launch build's.  Each thread keeps its elements
// of the current order in registers; the order a pass writes goes through
// global ping-pong buffers (L2-resident at these sizes), which
// __syncthreads orders.
#define SG_THREADS 1024
#define SG_WAVES (SG_THREADS / 64)
static_assert(PD_SMALL_GRID_MAX_M % (4 * SG_THREADS) == 0, "elements are handled four at a time");
static_assert(PD_SMALL_GRID_MAX_CELLS <= 65536, "at most two 8-bit passes");

struct SmallGridArgs {
  const double4* x;
  const int32_t* subset;
  int m;
  double mn[3];
  double cell;
  int c0[3];
  int dims[3];
  int ncell;
  int passes;
  uint32_t* keys_a;
  int32_t* idx_a;
  uint32_t* keys_b;
  int32_t* idx_b;
  int32_t* start;
  int32_t* end;
};

#define SG_PER (PD_SMALL_GRID_MAX_M / SG_THREADS)  // elements per thread

__global__ __launch_bounds__(SG_THREADS) void k_small_grid(SmallGridArgs a) {
  __shared__ uint32_t base[256];
  __shared__ uint32_t run[256];
  __shared__ uint32_t wcnt[SG_WAVES][256];
  const int t = threadIdx.x;
  const int w = t >> 6;
  const int lane = t & 63;
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  for (int c = t; c < a.ncell; c += SG_THREADS) {
    a.start[c] = 0;
    a.end[c] = 0;
  }
  // thread t holds elements j * SG_THREADS + t of the current order in
  // registers, so every chunk below is ranked without a load in its path
  uint32_t kr[SG_PER];
  int32_t vr[SG_PER];
#pragma unroll
  for (int h = 0; h < SG_PER; h += 4) {  // keys, four gathers in flight
    int pp[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = (h + j) * SG_THREADS + t;
      pp[j] = i < a.m ? a.subset[i] : -1;
    }
    double4 xx[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) xx[j] = pp[j] >= 0 ? a.x[pp[j]] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int cx, cy, cz;
      cell_of(xx[j], a.mn, a.cell, cx, cy, cz);
      cx -= a.c0[0];
      cy -= a.c0[1];
      cz -= a.c0[2];
      kr[h + j] = pp[j] >= 0 ? (uint32_t)(((int64_t)cz * a.dims[1] + cy) * a.dims[0] + cx) : 0u;
      vr[h + j] = pp[j];
    }
  }
  uint32_t* kin = a.keys_a;
  int32_t* vin = a.idx_a;
  uint32_t* kout = a.keys_b;
  int32_t* vout = a.idx_b;
  for (int ps = 0; ps < a.passes; ++ps) {
    const int shift = 8 * ps;
    if (ps > 0) {
#pragma unroll
      for (int j = 0; j < SG_PER; ++j) {
        const int i = j * SG_THREADS + t;
        kr[j] = i < a.m ? kin[i] : 0u;
        vr[j] = i < a.m ? vin[i] : 0;
      }
    }
    if (t < 256) {
      base[t] = 0;
      run[t] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SG_PER; ++j)
      if (j * SG_THREADS + t < a.m) atomicAdd(&base[(kr[j] >> shift) & 255u], 1u);
    __syncthreads();
    if (w == 0) {  // exclusive scan of the 256 digit counts, four per lane
      const uint32_t h0 = base[4 * lane], h1 = base[4 * lane + 1], h2 = base[4 * lane + 2],
                     h3 = base[4 * lane + 3];
      const uint32_t sum = h0 + h1 + h2 + h3;
      uint32_t incl = sum;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d, 64);
        if (lane >= d) incl += v;
      }
      const uint32_t ex = incl - sum;
      base[4 * lane] = ex;
      base[4 * lane + 1] = ex + h0;
      base[4 * lane + 2] = ex + h0 + h1;
      base[4 * lane + 3] = ex + h0 + h1 + h2;
    }
#pragma unroll
    for (int j = 0; j < SG_PER; ++j) {
      if (j * SG_THREADS >= a.m) break;  // uniform
      const bool valid = j * SG_THREADS + t < a.m;
      const uint32_t k = kr[j];
      const uint32_t d = (k >> shift) & 255u;
      for (int q = t; q < SG_WAVES * 256; q += SG_THREADS) (&wcnt[0][0])[q] = 0u;
      __syncthreads();
      unsigned long long peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long msk = __ballot(bit);
        peers &= bit ? msk : ~msk;
      }
      const uint32_t rank = (uint32_t)__popcll(peers & lt);
      if (valid && rank == 0) wcnt[w][d] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (valid) {
        uint32_t pos = base[d] + run[d] + rank;
        for (int ww = 0; ww < w; ++ww) pos += wcnt[ww][d];
        kout[pos] = k;
        vout[pos] = vr[j];
      }
      __syncthreads();
      if (t < 256) {
        uint32_t add = 0;
        for (int ww = 0; ww < SG_WAVES; ++ww) add += wcnt[ww][t];
        run[t] += add;
      }
      __syncthreads();  // wcnt is cleared next; run and kout are read next
    }
    uint32_t* tk = kin;
    kin = kout;
    kout = tk;
    int32_t* tv = vin;
    vin = vout;
    vout = tv;
  }
#pragma unroll
  for (int h = 0; h < SG_PER; h += 4) {  // cell ranges, the loads of four elements in flight
    uint32_t kc[4], kp[4], kn[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = (h + j) * SG_THREADS + t;
      kc[j] = i < a.m ? kin[i] : 0u;
      kp[j] = i > 0 && i < a.m ? kin[i - 1] : 0u;
      kn[j] = i + 1 < a.m ? kin[i + 1] : 0u;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int i = (h + j) * SG_THREADS + t;
      if (i < a.m) {
        if (i == 0 || kp[j] != kc[j]) a.start[kc[j]] = i;
        if (i == a.m - 1 || kn[j] != kc[j]) a.end[kc[j]] = i + 1;
      }
    }
  }
}

int launch_small_grid(const double4* x, const int32_t* subset, int32_t m, const double* mn,
                      double cell, const int32_t* c0, const int32_t* dims, int32_t ncell,
                      int32_t bits, uint32_t* keys_a, int32_t* idx_a, uint32_t* keys_b,
                      int32_t* idx_b, int32_t* start, int32_t* end, hipStream_t s) {
  SmallGridArgs a;
  a.x = x;
  a.subset = subset;
  a.m = m;
  for (int k = 0; k < 3; ++k) {
    a.mn[k] = mn[k];
    a.c0[k] = c0[k];
    a.dims[k] = dims[k];
  }
  a.cell = cell;
  a.ncell = ncell;
  a.passes = (bits + 7) / 8;
  a.keys_a = keys_a;
  a.idx_a = idx_a;
  a.keys_b = keys_b;
  a.idx_b = idx_b;
  a.start = start;
  a.end = end;
  hipLaunchKernelGGL(k_small_grid, dim3(1), dim3(SG_THREADS), 0, s, a);
  return a.passes & 1;
}

// ---------------------------------------------------------- neighbour test
// core/kdtree.h:66-79 dist() with target p: dx = x_p - x_q
__device__ __forceinline__ bool in_horizon(const double4 a, const double4 b, double h) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  return r < h;
}

struct NbArgs {
  int n;
  const double4* x;
  const int32_t* body;
  const int32_t* act;  // INVISIBLE particles take no bonds
  const double* horizon;  // per body
  GridView g;
  int32_t* counts;
  const int64_t* slice_off;
  int32_t* col;
  int32_t* top;
  int32_t* cap;
};

__global__ __launch_bounds__(PD_BLOCK) void k_bond_count(NbArgs a) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const double4 xp = a.x[p];
  const int bp = a.body[p];
  const double h = a.horizon[bp];
  if (a.act[p] == PD_INVISIBLE) {
    a.counts[p] = 0;
    return;
  }
  int cx, cy, cz;
  cell_of(xp, a.g.mn, a.g.cell, cx, cy, cz);
  int count = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int ix = cx + dx, iy = cy + dy, iz = cz + dz;
        if (ix < 0 || iy < 0 || iz < 0 || ix >= a.g.dims[0] || iy >= a.g.dims[1] ||
            iz >= a.g.dims[2])
          continue;
        const int64_t cell = ((int64_t)iz * a.g.dims[1] + iy) * a.g.dims[0] + ix;
        const int e = a.g.cell_end[cell];
        for (int m = a.g.cell_start[cell]; m < e; ++m) {
          const int q = a.g.sorted_idx[m];
          if (q == p || a.body[q] != bp || a.act[q] == PD_INVISIBLE) continue;
          if (in_horizon(xp, a.x[q], h)) ++count;
        }
      }
  a.counts[p] = count;
}

void launch_bond_count(int32_t n, const double4* x, const int32_t* body, const int32_t* act,
                       const double* horizon, const GridView& g, int32_t* counts, hipStream_t s) {
  NbArgs a = {};
  a.n = n;
  a.x = x;
  a.body = body;
  a.act = act;
  a.horizon = horizon;
  a.g = g;
  a.counts = counts;
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_bond_count, dim3(nb), dim3(PD_BLOCK), 0, s, a);
}

// max bond count of each 64-row slice
__global__ __launch_bounds__(PD_BLOCK) void k_slice_width(const int32_t* counts, int n,
                                                         int32_t* widths) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  int c = p < n ? counts[p] : 0;
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) c = max(c, __shfl_xor(c, m, 64));
  if ((threadIdx.x & 63) == 0 && p < n + 63) {
    const int s = p >> 6;
    if ((int64_t)s * 64 < n) widths[s] = c;
  }
}

void launch_slice_width(const int32_t* counts, int32_t n, int32_t* widths, hipStream_t s) {
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_slice_width, dim3(nb), dim3(PD_BLOCK), 0, s, counts, n, widths);
}

// One 64-thread workgroup per particle: the wave tests the 27 cells' members
// 64 at a time, compacts the hits into LDS with a ballot prefix, sorts them
// with a bitonic network and writes the row into its slice column.
#define FILL_CAP PD_MAX_LIST
#ifndef PD_MAX_LIST
#define PD_MAX_LIST 1024
#endif

__global__ __launch_bounds__(64) void k_bond_fill(NbArgs a) {
  __shared__ int list[PD_MAX_LIST];
  const int p = blockIdx.x;
  const int lane = threadIdx.x;
  const double4 xp = a.x[p];
  const int bp = a.body[p];
  const double h = a.horizon[bp];
  int cx, cy, cz;
  cell_of(xp, a.g.mn, a.g.cell, cx, cy, cz);
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  const bool visible = a.act[p] != PD_INVISIBLE;  // uniform over the workgroup
  int cnt = 0;
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        const int ix = cx + dx, iy = cy + dy, iz = cz + dz;
        if (ix < 0 || iy < 0 || iz < 0 || ix >= a.g.dims[0] || iy >= a.g.dims[1] ||
            iz >= a.g.dims[2])
          continue;
        const int64_t cell = ((int64_t)iz * a.g.dims[1] + iy) * a.g.dims[0] + ix;
        const int st = a.g.cell_start[cell];
        const int e = a.g.cell_end[cell];
        for (int m0 = st; m0 < e; m0 += 64) {
          const int m = m0 + lane;
          bool hit = false;
          int q = 0;
          if (visible && m < e) {
            q = a.g.sorted_idx[m];
            hit = (q != p) && (a.body[q] == bp) && (a.act[q] != PD_INVISIBLE) &&
                  in_horizon(xp, a.x[q], h);
          }
          const unsigned long long mask = __ballot(hit);
          if (hit) {
            const int pos = cnt + __popcll(mask & lt);
            if (pos < PD_MAX_LIST) list[pos] = q;
          }
          cnt += __popcll(mask);
        }
      }
  // the slice width was sized from k_bond_count's counts; never write past it
  if (cnt > a.counts[p]) cnt = a.counts[p];
  if (cnt > PD_MAX_LIST) cnt = PD_MAX_LIST;
  int P = 1;
  while (P < cnt) P <<= 1;
  for (int i = cnt + lane; i < P; i += 64) list[i] = 0x7fffffff;
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int u = list[i], v = list[ixj];
          const bool asc = (i & k) == 0;
          if ((u > v) == asc) {
            list[i] = v;
            list[ixj] = u;
          }
        }
      }
      __syncthreads();
    }
  }
  const int s = p >> 6;
  const int pl = p & 63;
  const int64_t base = a.slice_off[s];
  for (int k = lane; k < cnt; k += 64) a.col[(base + k) * 64 + pl] = list[k];
  if (lane == 0) {
    a.top[p] = cnt - 1;
    a.cap[p] = cnt;
  }
}

void launch_bond_fill(int32_t n, const double4* x, const int32_t* body, const int32_t* act,
                      const double* horizon, const GridView& g, const int32_t* counts,
                      const int64_t* slice_off, int32_t* col, int32_t* top, int32_t* cap,
                      hipStream_t s) {
  NbArgs a = {};
  a.counts = const_cast<int32_t*>(counts);
  a.n = n;
  a.x = x;
  a.body = body;
  a.act = act;
  a.horizon = horizon;
  a.g = g;
  a.slice_off = slice_off;
  a.col = col;
  a.top = top;
  a.cap = cap;
  if (n > 0) hipLaunchKernelGGL(k_bond_fill, dim3(n), dim3(64), 0, s, a);
}

}  // namespace pd
