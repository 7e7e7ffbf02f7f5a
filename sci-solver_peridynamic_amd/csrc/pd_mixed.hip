// pd_mixed.hip -- mixed-precision solve: iterative refinement in FP64
// around a block-Jacobi PCG on float vectors and float matrix blocks.
//
// The reference solves in FP64 only (core/simulator.cu:735-799); this solver
// keeps its system and stopping rule (||S(b - A x)|| <= rtol ||S b||, with
// the residual, its norm and every dot product in FP64) and moves the
// bandwidth-bound inner iterations to 4-byte values and vectors:
//   residual  r = S(b - A x) of the FP64 system: every FP64 bond block
//             recomputed from the positions with the assembly's arithmetic
//             (no FP64 blocks are stored), FP64 diagonal and x (bond terms
//             summed first in list order, the diagonal term added last);
//             fused with the inner start d = 0, z = M32 r32, p = z
//             (r32 = float(r)).  Refinement therefore converges to the
//             FP64 system's solution; A32 only serves the inner solves
//   inner     y = A32 p in float arithmetic (diagonal first, bonds in list
//             order, (a0 x0 + a1 x1) + a2 x2 per block row), alpha and beta
//             in FP64 from FP64 sums of widened float products, rounded to
//             float for the updates d += a p, r -= a y, p = z + b p
//   stop      inner: ||r32|| <= max(inner_rtol ||r||, tol / 2) or its
//             iteration limit; outer: ||r|| <= tol, refinement or total
//             iteration limit
//   correct   x += d in FP64
// Sums use the tree of the FP64 solver (wave butterfly, (w0+w1)+(w2+w3) per
// 256-row tile, strided fold of the tile partials), so the result is
// reproducible run to run and equal bit for bit to the oracle's restatement.
#include <math.h>

#include "pd_bond.h"

namespace pd {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));
typedef float f2v __attribute__((ext_vector_type(2)));

// Float bond blocks: slot k's 384 values hold a plane of (a00 a01 a02 a11)
// per lane, then a plane of (a12 a22) per lane, so a lane reads or writes
// its block with one 16-byte and one 8-byte access.
__device__ __forceinline__ void load_block32(const float* __restrict__ slot, int lane,
                                             float v[6]) {
  const f4v a = __builtin_nontemporal_load((const f4v*)slot + lane);
  const f2v b = __builtin_nontemporal_load((const f2v*)(slot + 256) + lane);
  v[0] = a.x;
  v[1] = a.y;
  v[2] = a.z;
  v[3] = a.w;
  v[4] = b.x;
  v[5] = b.y;
}


__device__ __forceinline__ void blk6w(float a00, float a01, float a02, float a11, float a12,
                                      float a22, const double4 x, double& o0, double& o1,
                                      double& o2) {
  const double b00 = a00, b01 = a01, b02 = a02, b11 = a11, b12 = a12, b22 = a22;
  o0 = b00 * x.x + b01 * x.y + b02 * x.z;
  o1 = b01 * x.x + b11 * x.y + b12 * x.z;
  o2 = b02 * x.x + b12 * x.y + b22 * x.z;
}

__device__ __forceinline__ void blk6f(float a00, float a01, float a02, float a11, float a12,
                                      float a22, const float4 x, float& o0, float& o1,
                                      float& o2) {
  o0 = a00 * x.x + a01 * x.y + a02 * x.z;
  o1 = a01 * x.x + a11 * x.y + a12 * x.z;
  o2 = a02 * x.x + a12 * x.y + a22 * x.z;
}

// Row p of A32 x in FP64 arithmetic (FP64 diagonal, float bonds); U slots
// per trip.
template <int U>
__device__ __forceinline__ void row_wide(const MxArgs& A, const double4* __restrict__ x, int p,
                                         double& y0, double& y1, double& y2) {
  const int s = p >> 6;
  const int lane = p & 63;
  const double* dg = A.diag + (int64_t)s * 6 * 64 + lane;
  // the bond terms first, from 0.0 in list order; the diagonal term is
  // added last (the order the assembly can produce in the same pass)
  y0 = 0.0;
  y1 = 0.0;
  y2 = 0.0;
  const int64_t base = A.slice_off[s];
  const int top = A.top[p];
  const int* __restrict__ cp = A.col + base * 64 + lane;
  const float* __restrict__ vp = A.val32 + base * 6 * 64;
  int bnd = 0;
  for (; bnd + U - 1 <= top; bnd += U) {
    int q[U];
    float v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = ntl(cp + (int64_t)(bnd + u) * 64);
#pragma unroll
    for (int u = 0; u < U; ++u) load_block32(vp + (int64_t)(bnd + u) * 384, lane, v[u]);
    double4 xx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xx[u] = x[q[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double t0, t1, t2;
      blk6w(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], xx[u], t0, t1, t2);
      y0 += t0;
      y1 += t1;
      y2 += t2;
    }
  }
  for (; bnd <= top; ++bnd) {
    const int q0 = ntl(cp + (int64_t)bnd * 64);
    float v[6];
    load_block32(vp + (int64_t)bnd * 384, lane, v);
    double t0, t1, t2;
    blk6w(v[0], v[1], v[2], v[3], v[4], v[5], x[q0], t0, t1, t2);
    y0 += t0;
    y1 += t1;
    y2 += t2;
  }
  {
    const double4 xp = x[p];
    const double a00 = dg[0], a01 = dg[64], a02 = dg[128], a11 = dg[192], a12 = dg[256],
                 a22 = dg[320];
    y0 = (a00 * xp.x + a01 * xp.y + a02 * xp.z) + y0;
    y1 = (a01 * xp.x + a11 * xp.y + a12 * xp.z) + y1;
    y2 = (a02 * xp.x + a12 * xp.y + a22 * xp.z) + y2;
  }
}

// Row p of A32 x in float arithmetic; U slots per trip.
template <int U>
__device__ __forceinline__ void row_f32(const MxArgs& A, const float4* __restrict__ x, int p,
                                        float& y0, float& y1, float& y2) {
  const int s = p >> 6;
  const int lane = p & 63;
  const float* dg = A.diag32 + (int64_t)s * 6 * 64 + lane;
  blk6f(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], x[p], y0, y1, y2);
  const int64_t base = A.slice_off[s];
  const int top = A.top[p];
  const int* __restrict__ cp = A.col + base * 64 + lane;
  const float* __restrict__ vp = A.val32 + base * 6 * 64;
  int bnd = 0;
  for (; bnd + U - 1 <= top; bnd += U) {
    int q[U];
    float v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = ntl(cp + (int64_t)(bnd + u) * 64);
#pragma unroll
    for (int u = 0; u < U; ++u) load_block32(vp + (int64_t)(bnd + u) * 384, lane, v[u]);
    float4 xx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xx[u] = x[q[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      float t0, t1, t2;
      blk6f(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], xx[u], t0, t1, t2);
      y0 += t0;
      y1 += t1;
      y2 += t2;
    }
  }
  for (; bnd <= top; ++bnd) {
    const int q0 = ntl(cp + (int64_t)bnd * 64);
    float v[6];
    load_block32(vp + (int64_t)bnd * 384, lane, v);
    float t0, t1, t2;
    blk6f(v[0], v[1], v[2], v[3], v[4], v[5], x[q0], t0, t1, t2);
    y0 += t0;
    y1 += t1;
    y2 += t2;
  }
}

// One bond's float block: the FP64 block recomputed from the positions as
// the assembly forms it (bond_dir, bond_jac, the time-step scaling),
// rounded to float -- the values the inner solves multiply.
template <bool SHARED>
__device__ __forceinline__ void bond_block32(const double4 pos_p, const double4 xq, double d0,
                                             double cb, double dt2, float v[6]) {
  double e0, e1, e2, dij;
  bond_dir<SHARED, SHARED>(pos_p, xq, e0, e1, e2, dij);
  double F00, F01, F02, F11, F12, F22;
  bond_jac<SHARED>(e0, e1, e2, dij, d0, cb, F00, F01, F02, F11, F12, F22);
  v[0] = (float)((F00 * -1.0) * dt2);
  v[1] = (float)((F01 * -1.0) * dt2);
  v[2] = (float)((F02 * -1.0) * dt2);
  v[3] = (float)((F11 * -1.0) * dt2);
  v[4] = (float)((F12 * -1.0) * dt2);
  v[5] = (float)((F22 * -1.0) * dt2);
}

// One block into slot `slot` of vals (the layout load_block32 reads).
__device__ __forceinline__ void store_block32(float* vals, int64_t slot, int lane,
                                              const float v[6]) {
  float* vo = vals + slot * 6 * 64;
  __builtin_nontemporal_store(f4v{v[0], v[1], v[2], v[3]}, (f4v*)vo + lane);
  __builtin_nontemporal_store(f2v{v[4], v[5]}, (f2v*)(vo + 256) + lane);
}

// Row p of A32 x as row_f32 computes it (the same float products and sums
// in the same order), the row's float blocks formed on the way
// (bond_block32) and written to vals.  REC: neighbour positions from the
// record A64.rec; SHARED: shared reciprocals (range flag clear).  V slots
// per trip.
template <int V, bool REC, bool SHARED>
__device__ __forceinline__ void row_f32_store(const MxArgs& A, float* __restrict__ vals,
                                              const float4* __restrict__ x, int p, float& y0,
                                              float& y1, float& y2) {
  const int s = p >> 6;
  const int lane = p & 63;
  const float* dg = A.diag32 + (int64_t)s * 6 * 64 + lane;
  blk6f(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], x[p], y0, y1, y2);
  const int top = A.top[p];
  if (top < 0) return;
  const SpmvView& F = A.A64;
  const int64_t base = A.slice_off[s];
  const int* __restrict__ cp = A.col + base * 64 + lane;
  const double* __restrict__ lp = F.d0 + base * 64 + lane;
  const double4 pos_p = F.pos[p];
  const double cb = F.mat[F.body[p]].c_bond;
  const double dt2 = F.consts ? F.consts->dt2 : F.dt2;
  auto pos_of = [&](int q) -> double4 {
    if constexpr (REC) {
      const dbl2v* r = (const dbl2v*)(F.rec + 6 * (int64_t)q);
      const dbl2v r0 = r[0], r1 = r[1];
      return make_double4(r0.x, r0.y, r1.x, 0.0);
    } else {
      return F.pos[q];
    }
  };
  int bnd = 0;
  for (; bnd + V - 1 <= top; bnd += V) {
    int q[V];
    double l0[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      q[u] = ntl(cp + (int64_t)(bnd + u) * 64);
      l0[u] = ntl(lp + (int64_t)(bnd + u) * 64);
    }
    double4 xq[V];
#pragma unroll
    for (int u = 0; u < V; ++u) xq[u] = pos_of(q[u]);
    float t[V][3];
#pragma unroll
    for (int u = 0; u < V; ++u) {
      float v[6];
      bond_block32<SHARED>(pos_p, xq[u], l0[u], cb, dt2, v);
      store_block32(vals, base + bnd + u, lane, v);
      blk6f(v[0], v[1], v[2], v[3], v[4], v[5], x[q[u]], t[u][0], t[u][1], t[u][2]);
    }
#pragma unroll
    for (int u = 0; u < V; ++u) {
      y0 += t[u][0];
      y1 += t[u][1];
      y2 += t[u][2];
    }
  }
  for (; bnd <= top; ++bnd) {
    const int q = ntl(cp + (int64_t)bnd * 64);
    float v[6];
    bond_block32<SHARED>(pos_p, pos_of(q), ntl(lp + (int64_t)bnd * 64), cb, dt2, v);
    store_block32(vals, base + bnd, lane, v);
    float t0, t1, t2;
    blk6f(v[0], v[1], v[2], v[3], v[4], v[5], x[q], t0, t1, t2);
    y0 += t0;
    y1 += t1;
    y2 += t2;
  }
}

__device__ __forceinline__ void precond_f(const float* minv, int pc, int p, float r0, float r1,
                                          float r2, float& z0, float& z1, float& z2) {
  if (pc != 0) {  // block Jacobi (1) or the learned block (2)
    const float* m = minv + (int64_t)(p >> 6) * 6 * 64 + (p & 63);
    const float m00 = m[0], m01 = m[64], m02 = m[128], m11 = m[192], m12 = m[256],
                m22 = m[320];
    z0 = m00 * r0 + m01 * r1 + m02 * r2;
    z1 = m01 * r0 + m11 * r1 + m12 * r2;
    z2 = m02 * r0 + m12 * r1 + m22 * r2;
  } else {
    z0 = r0;
    z1 = r1;
    z2 = r2;
  }
}

__device__ __forceinline__ double term3d(double a0, double a1, double a2, double b0, double b1,
                                         double b2) {
  return (a0 * b0 + a1 * b1) + a2 * b2;
}

// FP64 product of float operands (each widened exactly)
__device__ __forceinline__ double term3f(float a0, float a1, float a2, float b0, float b1,
                                         float b2) {
  return ((double)a0 * (double)b0 + (double)a1 * (double)b1) + (double)a2 * (double)b2;
}

__device__ __forceinline__ bool resid_due(int phase) { return phase == 0 || phase == 2; }

__global__ void k_mx_setup(pd_mx_state* states, int nspec, double rtol, int max_iter,
                           double inner_rtol, int inner_max, int max_refine) {
  for (int k = threadIdx.x; k < nspec; k += blockDim.x) {
    pd_mx_state* st = states + k;
    st->rz = 0.0;
    st->rr32 = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->rnorm = 0.0;
    st->bnorm = 0.0;
    st->tol = 0.0;
    st->inner_tol = 0.0;
    st->rtol = rtol;
    st->inner_rtol = inner_rtol;
    st->iter = 0;
    st->inner_iter = 0;
    st->refinements = 0;
    st->phase = 0;
    st->converged = 0;
    st->first = 1;
    st->max_iter = max_iter;
    st->inner_max = inner_max;
    st->max_refine = max_refine;
    st->residuals = 0;
  }
}

// r = S(b - A x) in FP64 (the FP64 system, blocks recomputed); r32 =
// float(r); z = M32 r32; p = z; d = 0; partials <r,r>, <r32,z>, <r32,r32>,
// <Sb,Sb>.  After a mixed assembly (A64.rec set: the record holds the
// positions and x side by side) the neighbours come from the record, four
// bonds per trip with shared reciprocals when the range flag is clear, two
// with IEEE quotients when it is set; otherwise from the arrays.
__global__ __launch_bounds__(PD_BLOCK) void k_mx_resid(MxArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (!resid_due(a.st[spec_of(a.batch, tile)].phase)) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t_rr = 0.0, t_rz = 0.0, t_r32 = 0.0, t_bb = 0.0;
  if (p < a.n) {
    double y0, y1, y2;
    if (!a.A64.rec)
      spmv_row_mf<3, true, false, false>(a.A64, a.x, p, y0, y1, y2);
    else if (*a.A64.range_bad == 0)  // uniform over the grid
      spmv_row_mf<4, true, true, true>(a.A64, a.x, p, y0, y1, y2, a.A64.rec);
    else
      spmv_row_mf<2, true, true, false>(a.A64, a.x, p, y0, y1, y2, a.A64.rec);
    const double4 bp = a.b[p];
    const bool fr = a.bc[p] == 0;
    const double r0 = fr ? bp.x - y0 : 0.0;
    const double r1 = fr ? bp.y - y1 : 0.0;
    const double r2 = fr ? bp.z - y2 : 0.0;
    const double sb0 = fr ? bp.x : 0.0, sb1 = fr ? bp.y : 0.0, sb2 = fr ? bp.z : 0.0;
    const float f0 = (float)r0, f1 = (float)r1, f2 = (float)r2;
    float z0, z1, z2;
    precond_f(a.minv32, a.precond, p, f0, f1, f2, z0, z1, z2);
    a.r[p] = make_float4(f0, f1, f2, 0.0f);
    a.p[p] = make_float4(z0, z1, z2, 0.0f);
    a.d[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    t_rr = term3d(r0, r1, r2, r0, r1, r2);
    t_rz = term3f(f0, f1, f2, z0, z1, z2);
    t_r32 = term3f(f0, f1, f2, f0, f1, f2);
    t_bb = term3d(sb0, sb1, sb2, sb0, sb1, sb2);
  }
  const double v0 = tile_sum(t_rr, lds4);
  const double v1 = tile_sum(t_rz, lds4);
  const double v2 = tile_sum(t_r32, lds4);
  const double v3 = tile_sum(t_bb, lds4);
  if (threadIdx.x == 0) {
    a.part[tile] = v0;
    a.part[ntiles + tile] = v1;
    a.part[2 * (int64_t)ntiles + tile] = v2;
    a.part[3 * (int64_t)ntiles + tile] = v3;
  }
}

// outer test; on failure of the test start the next inner solve
// one workgroup per specimen
__global__ __launch_bounds__(PD_BLOCK) void k_mx_resid_fold(MxArgs a, int ntiles) {
  if (!resid_due(a.st[blockIdx.x].phase)) return;
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  const double rr = fold_partials(a.part + t0, nt, lds4);
  const double rz = fold_partials(a.part + ntiles + t0, nt, lds4);
  const double rr32 = fold_partials(a.part + 2 * (int64_t)ntiles + t0, nt, lds4);
  const double bb = fold_partials(a.part + 3 * (int64_t)ntiles + t0, nt, lds4);
  if (threadIdx.x == 0) {
    pd_mx_state* st = a.st + blockIdx.x;
    st->residuals = st->residuals + 1;
    if (st->first) {
      st->bnorm = sqrt(bb);
      st->tol = st->rtol * st->bnorm;
      st->first = 0;
    }
    const double rnorm = sqrt(rr);
    st->rnorm = rnorm;
    if (rnorm <= st->tol) {
      st->converged = 1;
      st->phase = 3;
    } else if (st->refinements >= st->max_refine || st->iter >= st->max_iter) {
      st->phase = 3;
    } else {
      st->refinements = st->refinements + 1;
      const double at = st->inner_rtol * rnorm, bt = 0.5 * st->tol;
      st->inner_tol = at > bt ? at : bt;
      st->rz = rz;
      st->rr32 = rr32;
      st->inner_iter = 0;
      const bool stop = sqrt(rr32) <= st->inner_tol || st->inner_max <= 0 ||
                        st->iter >= st->max_iter;
      st->phase = stop ? 2 : 1;
    }
  }
}

// y = A32 p (float); partial <y,p>
template <int U>
__global__ __launch_bounds__(PD_BLOCK) void k_mx_spmv(MxArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (a.st[spec_of(a.batch, tile)].phase != 1) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t = 0.0;
  if (p < a.n) {
    float y0, y1, y2;
    row_f32<U>(a, a.p, p, y0, y1, y2);
    a.y[p] = make_float4(y0, y1, y2, 0.0f);
    const float4 pp = a.p[p];
    t = term3f(y0, y1, y2, pp.x, pp.y, pp.z);
  }
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[tile] = v;
}

// The first inner product of a solve after a mixed assembly, which leaves
// the float blocks to it: y = A32 p and the <y,p> partial as k_mx_spmv
// (for the specimens in their inner loop), and every row's float blocks
// written to vals, converged specimens' too, so the step leaves the whole
// float matrix.
__global__ __launch_bounds__(PD_BLOCK) void k_mx_spmv_store(MxArgs a, float* vals) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const bool act = a.st[spec_of(a.batch, tile)].phase == 1;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t = 0.0;
  if (p < a.n) {
    float y0, y1, y2;
    if (!a.A64.rec)
      row_f32_store<3, false, false>(a, vals, a.p, p, y0, y1, y2);
    else if (*a.A64.range_bad == 0)  // uniform over the grid
      row_f32_store<3, true, true>(a, vals, a.p, p, y0, y1, y2);
    else
      row_f32_store<3, true, false>(a, vals, a.p, p, y0, y1, y2);
    if (act) {
      a.y[p] = make_float4(y0, y1, y2, 0.0f);
      const float4 pp = a.p[p];
      t = term3f(y0, y1, y2, pp.x, pp.y, pp.z);
    }
  }
  if (!act) return;  // uniform over the tile
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[tile] = v;
}

// The float blocks of the current bond lists from the positions (IEEE
// quotients), for callers that read them without a solve having run.
__global__ __launch_bounds__(PD_BLOCK) void k_write_blocks32(SpmvView A, float* vals) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= A.n) return;
  const int top = A.top[p];
  if (top < 0) return;
  const int s = p >> 6, lane = p & 63;
  const int64_t base = A.slice_off[s];
  const double4 pos_p = A.pos[p];
  const double cb = A.mat[A.body[p]].c_bond;
  const double dt2 = A.consts ? A.consts->dt2 : A.dt2;
  for (int bnd = 0; bnd <= top; ++bnd) {
    const int64_t i = (base + bnd) * 64 + lane;
    float v[6];
    bond_block32<false>(pos_p, A.pos[A.col[i]], A.d0[i], cb, dt2, v);
    store_block32(vals, base + bnd, lane, v);
  }
}

// alpha = rz / <y,p>, formed by every workgroup from its specimen's <y,p>
// partials in the one fold order; then d += a p; r -= a y (free rows);
// z = M32 r; partials <r,z>, <r,r> into the third and fourth partial rows
// (the first row is still being folded by other workgroups)
__global__ __launch_bounds__(PD_BLOCK) void k_mx_update(MxArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int s = spec_of(a.batch, tile);
  pd_mx_state* st = a.st + s;
  if (st->phase != 1) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  const bool row = p < a.n;
  // the row's loads go out ahead of the fold
  float4 pp = make_float4(0.0f, 0.0f, 0.0f, 0.0f), dd = pp, rr = pp, yy = pp;
  bool fr = false;
  if (row) {
    pp = a.p[p];
    dd = a.d[p];
    rr = a.r[p];
    fr = a.bc[p] == 0;
    if (fr) yy = a.y[p];
  }
  int t0, nt;
  spec_range_of(a.batch, s, ntiles, t0, nt);
  double alpha;
  if (nt > PD_ALPHA_FOLD_TILES) {
    alpha = st->alpha;  // k_mx_alpha folded it ahead of this launch
  } else {
    const double yp = fold_partials(a.part + t0, nt, lds4);
    alpha = st->rz / yp;
    if (tile == t0 && threadIdx.x == 0) st->alpha = alpha;
  }
  const float af = (float)alpha;
  double t_rz = 0.0, t_rr = 0.0;
  if (row) {
    dd.x = af * pp.x + dd.x;
    dd.y = af * pp.y + dd.y;
    dd.z = af * pp.z + dd.z;
    a.d[p] = dd;
    if (fr) {
      const float na = -af;
      rr.x = na * yy.x + rr.x;
      rr.y = na * yy.y + rr.y;
      rr.z = na * yy.z + rr.z;
      a.r[p] = rr;
    }
    float z0, z1, z2;
    precond_f(a.minv32, a.precond, p, rr.x, rr.y, rr.z, z0, z1, z2);
    a.z[p] = make_float4(z0, z1, z2, 0.0f);
    t_rz = term3f(rr.x, rr.y, rr.z, z0, z1, z2);
    t_rr = term3f(rr.x, rr.y, rr.z, rr.x, rr.y, rr.z);
  }
  const double v0 = tile_sum(t_rz, lds4);
  const double v1 = tile_sum(t_rr, lds4);
  if (threadIdx.x == 0) {
    a.part[2 * (int64_t)ntiles + tile] = v0;
    a.part[3 * (int64_t)ntiles + tile] = v1;
  }
}

// alpha of the specimens with more than PD_ALPHA_FOLD_TILES tiles, one
// workgroup each (the same fold and quotient as k_mx_update's)
__global__ __launch_bounds__(PD_BLOCK) void k_mx_alpha(MxArgs a, int ntiles) {
  pd_mx_state* st = a.st + blockIdx.x;
  if (st->phase != 1) return;
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  if (nt <= PD_ALPHA_FOLD_TILES) return;
  const double yp = fold_partials(a.part + t0, nt, lds4);
  if (threadIdx.x == 0) st->alpha = st->rz / yp;
}

// beta and the inner test ahead of the next inner iteration
__global__ __launch_bounds__(PD_BLOCK) void k_mx_beta(MxArgs a, int ntiles) {
  if (a.st[blockIdx.x].phase != 1) return;
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  const double rz = fold_partials(a.part + 2 * (int64_t)ntiles + t0, nt, lds4);
  const double rr32 = fold_partials(a.part + 3 * (int64_t)ntiles + t0, nt, lds4);
  if (threadIdx.x == 0) {
    pd_mx_state* st = a.st + blockIdx.x;
    st->beta = rz / st->rz;
    st->rz = rz;
    st->rr32 = rr32;
    st->iter = st->iter + 1;
    st->inner_iter = st->inner_iter + 1;
    if (sqrt(rr32) <= st->inner_tol || st->inner_iter >= st->inner_max ||
        st->iter >= st->max_iter)
      st->phase = 2;
  }
}

// p = z + b p
__global__ __launch_bounds__(PD_BLOCK) void k_mx_dir(MxArgs a) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const pd_mx_state* st = a.st + spec_of(a.batch, tile);
  if (st->phase != 1) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const float bf = (float)st->beta;
  const float4 zz = a.z[p];
  float4 pp = a.p[p];
  pp.x = zz.x + bf * pp.x;
  pp.y = zz.y + bf * pp.y;
  pp.z = zz.z + bf * pp.z;
  a.p[p] = pp;
}

// x += d (FP64), and x into the record (A64.rec set) for the residual's
// neighbour gathers
__global__ __launch_bounds__(PD_BLOCK) void k_mx_apply(MxArgs a) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (a.st[spec_of(a.batch, tile)].phase != 2) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const float4 dd = a.d[p];
  double4 xx = a.x[p];
  xx.x = xx.x + (double)dd.x;
  xx.y = xx.y + (double)dd.y;
  xx.z = xx.z + (double)dd.z;
  a.x[p] = xx;
  if (a.A64.rec) {
    double* rp = a.A64.rec + 6 * (int64_t)p;
    rp[3] = xx.x;
    rp[4] = xx.y;
    rp[5] = xx.z;
  }
}

// x into the record (A64.rec set) ahead of a residual no apply preceded:
// the rows of specimens whose residual is due
__global__ __launch_bounds__(PD_BLOCK) void k_mx_rec_x(MxArgs a) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  if (!resid_due(a.st[spec_of(a.batch, tile)].phase)) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  const double4 xx = a.x[p];
  double* rp = a.A64.rec + 6 * (int64_t)p;
  rp[3] = xx.x;
  rp[4] = xx.y;
  rp[5] = xx.z;
}

template <int U>
__global__ __launch_bounds__(PD_BLOCK) void k_spmv_wide(MxArgs a, const double4* __restrict__ x,
                                                       double4* __restrict__ y) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  double y0, y1, y2;
  row_wide<U>(a, x, p, y0, y1, y2);
  y[p] = make_double4(y0, y1, y2, 0.0);
}

template <int U>
__global__ __launch_bounds__(PD_BLOCK) void k_spmv_f32(MxArgs a, const float4* __restrict__ x,
                                                      float4* __restrict__ y) {
  const int p = xcd_tile(blockIdx.x, gridDim.x) * PD_BLOCK + threadIdx.x;
  if (p >= a.n) return;
  float y0, y1, y2;
  row_f32<U>(a, x, p, y0, y1, y2);
  y[p] = make_float4(y0, y1, y2, 0.0f);
}

inline int tiles_of(int n) { return (n + PD_BLOCK - 1) / PD_BLOCK; }

}  // namespace

void launch_mx_setup(pd_mx_state* st, int32_t nspec, double rtol, int32_t max_iter,
                     double inner_rtol, int32_t inner_max, int32_t max_refine, hipStream_t s) {
  hipLaunchKernelGGL(k_mx_setup, dim3(1), dim3(256), 0, s, st, nspec, rtol, max_iter, inner_rtol,
                     inner_max, max_refine);
}
void launch_mx_rec_x(const MxArgs& a, hipStream_t s) {
  if (a.n && a.A64.rec)
    hipLaunchKernelGGL(k_mx_rec_x, dim3(tiles_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_mx_resid(const MxArgs& a, hipStream_t s) {
  if (!a.n) return;
  const int g = tiles_of(a.n);
  hipLaunchKernelGGL(k_mx_resid, dim3(g), dim3(PD_BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mx_resid_fold, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g);
}
void launch_mx_resid_fold(const MxArgs& a, hipStream_t s) {
  if (a.n)
    hipLaunchKernelGGL(k_mx_resid_fold, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a,
                       tiles_of(a.n));
}
void launch_mx_spmv(const MxArgs& a, hipStream_t s) {
  if (!a.n) return;
  const dim3 g(tiles_of(a.n));
  switch (a.unroll) {
    case 2: hipLaunchKernelGGL(k_mx_spmv<2>, g, dim3(PD_BLOCK), 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_mx_spmv<8>, g, dim3(PD_BLOCK), 0, s, a); break;
    default: hipLaunchKernelGGL(k_mx_spmv<4>, g, dim3(PD_BLOCK), 0, s, a);
  }
}
void launch_mx_spmv_store(const MxArgs& a, float* vals, hipStream_t s) {
  if (a.n)
    hipLaunchKernelGGL(k_mx_spmv_store, dim3(tiles_of(a.n)), dim3(PD_BLOCK), 0, s, a, vals);
}
void launch_write_blocks32(const SpmvView& A, float* vals, hipStream_t s) {
  if (A.n) hipLaunchKernelGGL(k_write_blocks32, dim3(tiles_of(A.n)), dim3(PD_BLOCK), 0, s, A, vals);
}
void launch_mx_update(const MxArgs& a, hipStream_t s) {
  if (!a.n) return;
  const int g = tiles_of(a.n);
  if (g > PD_ALPHA_FOLD_TILES)  // some specimen may exceed the threshold
    hipLaunchKernelGGL(k_mx_alpha, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g);
  hipLaunchKernelGGL(k_mx_update, dim3(g), dim3(PD_BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_mx_beta, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g);
}
void launch_mx_dir(const MxArgs& a, hipStream_t s) {
  if (a.n) hipLaunchKernelGGL(k_mx_dir, dim3(tiles_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_mx_apply(const MxArgs& a, hipStream_t s) {
  if (a.n) hipLaunchKernelGGL(k_mx_apply, dim3(tiles_of(a.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_spmv_wide(const MxArgs& a, const double4* x, double4* y, hipStream_t s) {
  if (!a.n) return;
  const dim3 g(tiles_of(a.n));
  switch (a.unroll) {
    case 2: hipLaunchKernelGGL(k_spmv_wide<2>, g, dim3(PD_BLOCK), 0, s, a, x, y); break;
    case 8: hipLaunchKernelGGL(k_spmv_wide<8>, g, dim3(PD_BLOCK), 0, s, a, x, y); break;
    default: hipLaunchKernelGGL(k_spmv_wide<4>, g, dim3(PD_BLOCK), 0, s, a, x, y);
  }
}
void launch_spmv_f32(const MxArgs& a, const float4* x, float4* y, hipStream_t s) {
  if (!a.n) return;
  const dim3 g(tiles_of(a.n));
  switch (a.unroll) {
    case 2: hipLaunchKernelGGL(k_spmv_f32<2>, g, dim3(PD_BLOCK), 0, s, a, x, y); break;
    case 8: hipLaunchKernelGGL(k_spmv_f32<8>, g, dim3(PD_BLOCK), 0, s, a, x, y); break;
    default: hipLaunchKernelGGL(k_spmv_f32<4>, g, dim3(PD_BLOCK), 0, s, a, x, y);
  }
}

}  // namespace pd
