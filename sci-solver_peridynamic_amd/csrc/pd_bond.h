// pd_bond.h -- per-bond FP64 arithmetic shared by the assembly, the
// matrix-free SpMV and the products that write the bond blocks, and the
// streaming-load helpers of the kernels (included by the HIP sources only).
//
// Floating point follows the reference's host semantics bit for bit: every
// including file is compiled with -ffp-contract=off (no FMA contraction),
// IEEE division and square root, and the two float roundings of
// cg_solver.cuh:256,264.
#pragma once

#include <math.h>

#include "pd_internal.h"

namespace pd {

__device__ __forceinline__ double ntl(const double* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ float ntl(const float* p) { return __builtin_nontemporal_load(p); }
__device__ __forceinline__ int ntl(const int* p) { return __builtin_nontemporal_load(p); }

// two FP64 block values of one lane (the pair planes of val)
typedef double vpair __attribute__((ext_vector_type(2)));

// Division by a reciprocal shared between quotients.  gfx950 divides a / b
// in FP64 as: bs = div_scale(b), as = div_scale(a); y = rcp(bs) refined by
// two Newton steps y += y (1 - bs y); q = as y; the Markstein correction
// fma(as - bs q, y, q) (div_fmas); div_fixup for special operands and the
// sign.  When no operand needs scaling (div_scale leaves both unchanged and
// div_fmas applies no 2^64 factor) and no special case applies, that is
// exactly rcp_nr(b) then div_rcp(a, b, y) below, so quotients by one
// denominator can share y.  No scaling happens when b and 1/b are normal,
// a is 0 or |a| >= 2^-969, a / b is normal and exponent(a) - exponent(b) <
// 768.  Every denominator here is positive: then a = +0 gives +0 as IEEE
// division does, but a = -0 would give +0 where it gives -0, so no caller
// passes -0 (a coordinate difference is -0 only for a -0 coordinate less a
// +0 one, and the range flag admits no -0 coordinate; the other numerators
// are positive or a difference horizon - d0 with a positive horizon, which
// is never -0).  range_ok_* state sufficient ranges; callers take this path
// only when the assembly found every operand inside them.
__device__ __forceinline__ double rcp_nr(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double t = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, t, y);
  t = __builtin_fma(-b, y, 1.0);
  y = __builtin_fma(y, t, y);
  return y;
}
__device__ __forceinline__ double div_rcp(double a, double b, double y) {
  const double q = a * y;
  const double r = __builtin_fma(-b, q, a);
  return __builtin_fma(r, y, q);
}
// Square root of x in [2^-767, 2^1023] as gfx950 forms IEEE sqrt() for
// such x, without the input scaling sqrt() applies below 2^-767 and the
// selects for zero and infinity: g = rsq(x); s = x g, h = g / 2 refined
// together by one coupled Newton step, then two Markstein corrections
// s += (x - s^2) h.  The same operations in the same order, so the same
// bits.  x = 0 gives NaN: callers exclude it.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double g = __builtin_amdgcn_rsq(x);
  double s = x * g;
  double h = g * 0.5;
  const double r = __builtin_fma(-h, s, 0.5);
  s = __builtin_fma(s, r, s);
  h = __builtin_fma(h, r, h);
  double d = __builtin_fma(-s, s, x);
  s = __builtin_fma(d, h, s);
  d = __builtin_fma(-s, s, x);
  return __builtin_fma(d, h, s);
}

// A coordinate +0 or of magnitude in [2^-300, 2^100]: a difference of two
// such values is +0 or of magnitude in [2^-352, 2^101], so
//  * it divided by the bond length fd = max(|d|, 1e-10) (at most 2^102, a
//    float) needs no scaling;
//  * the squared length of a bond is 0 or in [2^-704, 2^204], inside
//    sqrt_nr's range when it is not 0, and the length dij is 0 or in
//    [2^-352, 2^102].
__device__ __forceinline__ bool range_ok_coord(double v) {
  const double m = fabs(v);
  return (m == 0.0 && !__builtin_signbit(v)) || (m >= 0x1p-300 && m <= 0x1p100);
}
// With d0 in [2^-400, 2^100] and c_bond, the horizon and the radius in
// [2^-300, 2^300]: c_bond / d0; dij / d0 in [2^-502, 2^502]; d0 / dij for a
// nonzero dij in [2^-502, 2^452]; and (horizon - d0) / 2 / radius, whose
// numerator is 0 or at least ulp(d0) >= 2^-452 in magnitude, need no
// scaling either.
__device__ __forceinline__ bool range_ok_len(double d0) {
  return d0 >= 0x1p-400 && d0 <= 0x1p100;
}
__device__ __forceinline__ bool range_ok_stiff(double c) {
  return c >= 0x1p-300 && c <= 0x1p300;
}

// Current geometry of bond (p, q), cg_solver.cuh:253-256: the unit vector e
// (the difference divided by the float-rounded length max(|d|, SMALL)) and
// the current length dij.  SHARED: the three quotients share one
// reciprocal, and the square root is sqrt_nr, behind a test for a zero
// length unless NONZERO (the caller knows every bond length is nonzero):
// identical results inside the ranges above.
template <bool SHARED = false, bool NONZERO = false>
__device__ __forceinline__ void bond_dir(const double4 pos_p, const double4 xq, double& e0,
                                         double& e1, double& e2, double& dij) {
  e0 = xq.x - pos_p.x;
  e1 = xq.y - pos_p.y;
  e2 = xq.z - pos_p.z;
  const double ss = e0 * e0 + e1 * e1 + e2 * e2;
  if constexpr (!SHARED)
    dij = sqrt(ss);
  else if constexpr (NONZERO)
    dij = sqrt_nr(ss);
  else
    dij = ss == 0.0 ? 0.0 : sqrt_nr(ss);
  const float dij_f = (float)dij;
  const float small_f = (float)1.0e-10;
  const double fd = (double)(dij_f > small_f ? dij_f : small_f);
  if constexpr (SHARED) {
    const double y = rcp_nr(fd);
    e0 = div_rcp(e0, fd, y);
    e1 = div_rcp(e1, fd, y);
    e2 = div_rcp(e2, fd, y);
  } else {
    e0 /= fd;
    e1 /= fd;
    e2 /= fd;
  }
}

// Reference length of bond (p, q), cg_solver.cuh:258
__device__ __forceinline__ double ref_len(const double4 opos_p, const double4 oq) {
  const double o0 = opos_p.x - oq.x, o1 = opos_p.y - oq.y, o2 = opos_p.z - oq.z;
  return sqrt(o0 * o0 + o1 * o1 + o2 * o2);
}

// The bond's 3x3 Jacobian F (six distinct entries), cg_solver.cuh:292-299:
// (c / d0) [(1 - d0/dij) I + (d0/dij) e e^T], with the factor sc = c / d0
// given.  SHARED: d0 / dij through rcp_nr(dij) (identical inside the ranges
// above for a nonzero dij; a zero dij makes every entry NaN either way).
template <bool SHARED = false>
__device__ __forceinline__ void bond_jac_sc(double e0, double e1, double e2, double dij,
                                            double d0, double sc, double& F00, double& F01,
                                            double& F02, double& F11, double& F12,
                                            double& F22) {
  const double ratio = SHARED ? div_rcp(d0, dij, rcp_nr(dij)) : d0 / dij;
  const double dg = 1.0 - ratio;
  F00 = (dg + (e0 * e0) * ratio) * sc;
  F01 = (0.0 + (e0 * e1) * ratio) * sc;
  F02 = (0.0 + (e0 * e2) * ratio) * sc;
  F11 = (dg + (e1 * e1) * ratio) * sc;
  F12 = (0.0 + (e1 * e2) * ratio) * sc;
  F22 = (dg + (e2 * e2) * ratio) * sc;
}

// The same from c: SHARED forms c / d0 through rcp_nr (identical results
// inside the ranges above).
template <bool SHARED = false>
__device__ __forceinline__ void bond_jac(double e0, double e1, double e2, double dij, double d0,
                                         double c_bond, double& F00, double& F01, double& F02,
                                         double& F11, double& F12, double& F22) {
  double sc;
  if constexpr (SHARED)
    sc = div_rcp(c_bond, d0, rcp_nr(d0));
  else
    sc = c_bond / d0;
  bond_jac_sc<SHARED>(e0, e1, e2, dij, d0, sc, F00, F01, F02, F11, F12, F22);
}

__device__ __forceinline__ void blk6(double a00, double a01, double a02, double a11,
                                     double a12, double a22, const double4 x,
                                     double& o0, double& o1, double& o2) {
  o0 = a00 * x.x + a01 * x.y + a02 * x.z;
  o1 = a01 * x.x + a11 * x.y + a12 * x.z;
  o2 = a02 * x.x + a12 * x.y + a22 * x.z;
}


// Row p of A x, matrix-free: each live bond's block is recomputed from the
// step's positions (bond_dir), the slot's stored reference length and
// bond_jac, and scaled as the assembly scales it, so the blocks, and the
// products, equal the stored ones bit for bit (same summation order as
// spmv_row).  Bound by the vector memory pipeline (gathers) and FP64
// arithmetic, not by HBM.  REC: the neighbour's position and x come from
// A.rec (x must be the CG direction p the record holds): three 16-byte
// loads instead of four.  SHARED: quotients share reciprocals and square
// roots take sqrt_nr with no zero test (the assembly found every operand in
// range and no live bond of zero length).
typedef double dbl2v __attribute__((ext_vector_type(2)));

// One bond block into the pair planes of vals (the assembly's layout).
__device__ __forceinline__ void store_block(double* vals, int64_t slot, int lane, double b00,
                                            double b01, double b02, double b11, double b12,
                                            double b22) {
  vpair* vo = (vpair*)(vals + slot * 6 * 64) + lane;
  __builtin_nontemporal_store(vpair{b00, b01}, vo);
  __builtin_nontemporal_store(vpair{b02, b11}, vo + 64);
  __builtin_nontemporal_store(vpair{b12, b22}, vo + 128);
}

// The second half of particle q's record (the vector the record carries).
__device__ __forceinline__ double4 rec_vec(const double* __restrict__ rec, int q) {
  const dbl2v* r = (const dbl2v*)(rec + 6 * (int64_t)q);
  const dbl2v r1 = r[1], r2 = r[2];
  return make_double4(r1.y, r2.x, r2.y, 0.0);
}

template <bool REC>
__device__ __forceinline__ void gather_mf(const SpmvView& A, const double* __restrict__ rec,
                                          const double4* __restrict__ x, int q, double4& xq,
                                          double4& vq) {
  if constexpr (REC) {
    const dbl2v* r = (const dbl2v*)(rec + 6 * (int64_t)q);
    const dbl2v r0 = r[0], r1 = r[1], r2 = r[2];
    xq = make_double4(r0.x, r0.y, r1.x, 0.0);
    vq = make_double4(r1.y, r2.x, r2.y, 0.0);
  } else {
    xq = A.pos[q];
    vq = x[q];
  }
}

// STORE: each recomputed block is also written to vals (pair planes).
template <int V, bool DIAG_LAST, bool REC, bool SHARED, bool STORE = false>
__device__ __forceinline__ void spmv_row_mf(const SpmvView& A, const double4* __restrict__ x,
                                            int p, double& y0, double& y1, double& y2,
                                            const double* __restrict__ rec = nullptr,
                                            double* __restrict__ vals = nullptr) {
  const int s = p >> 6;
  const int lane = p & 63;
  const double* dg = A.diag + (int64_t)s * 6 * 64 + lane;
  // REC: the row's own value from its record too (x may not hold it)
  if constexpr (DIAG_LAST) {
    y0 = 0.0;
    y1 = 0.0;
    y2 = 0.0;
  } else {
    blk6(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], REC ? rec_vec(rec, p) : x[p], y0,
         y1, y2);
  }
  const int top = A.top[p];
  if (top >= 0) {
    const int64_t base = A.slice_off[s];
    const int* __restrict__ cp = A.col + base * 64 + lane;
    const double4 pos_p = A.pos[p];
    const double cb = A.mat[A.body[p]].c_bond;
    const double dt2 = A.consts ? A.consts->dt2 : A.dt2;
    const double* __restrict__ lp = A.d0 + base * 64 + lane;
    // V slots per trip: their blocks are independent chains of FP64
    // arithmetic (divisions, square roots) the scheduler interleaves
    int bnd = 0;
    for (; bnd + V - 1 <= top; bnd += V) {
      int q[V];
      double l0[V];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        q[u] = ntl(cp + (int64_t)(bnd + u) * 64);
        l0[u] = ntl(lp + (int64_t)(bnd + u) * 64);
      }
      double4 xq[V], vq[V];
#pragma unroll
      for (int u = 0; u < V; ++u) gather_mf<REC>(A, rec, x, q[u], xq[u], vq[u]);
      double t[V][3];
#pragma unroll
      for (int u = 0; u < V; ++u) {
        double e0, e1, e2, dij;
        bond_dir<SHARED, SHARED>(pos_p, xq[u], e0, e1, e2, dij);
        double F00, F01, F02, F11, F12, F22;
        bond_jac<SHARED>(e0, e1, e2, dij, l0[u], cb, F00, F01, F02, F11, F12, F22);
        const double b00 = (F00 * -1.0) * dt2, b01 = (F01 * -1.0) * dt2,
                     b02 = (F02 * -1.0) * dt2, b11 = (F11 * -1.0) * dt2,
                     b12 = (F12 * -1.0) * dt2, b22 = (F22 * -1.0) * dt2;
        if constexpr (STORE) store_block(vals, base + bnd + u, lane, b00, b01, b02, b11, b12, b22);
        blk6(b00, b01, b02, b11, b12, b22, vq[u], t[u][0], t[u][1], t[u][2]);
      }
#pragma unroll
      for (int u = 0; u < V; ++u) {
        y0 += t[u][0];
        y1 += t[u][1];
        y2 += t[u][2];
      }
    }
    for (; bnd <= top; ++bnd) {
      const int q0 = ntl(cp + (int64_t)bnd * 64);
      double4 xq, vq;
      gather_mf<REC>(A, rec, x, q0, xq, vq);
      double e0, e1, e2, dij;
      bond_dir<SHARED, SHARED>(pos_p, xq, e0, e1, e2, dij);
      double F00, F01, F02, F11, F12, F22;
      bond_jac<SHARED>(e0, e1, e2, dij, ntl(lp + (int64_t)bnd * 64), cb, F00, F01, F02, F11,
                       F12, F22);
      const double b00 = (F00 * -1.0) * dt2, b01 = (F01 * -1.0) * dt2, b02 = (F02 * -1.0) * dt2,
                   b11 = (F11 * -1.0) * dt2, b12 = (F12 * -1.0) * dt2, b22 = (F22 * -1.0) * dt2;
      if constexpr (STORE) store_block(vals, base + bnd, lane, b00, b01, b02, b11, b12, b22);
      double t0, t1, t2;
      blk6(b00, b01, b02, b11, b12, b22, vq, t0, t1, t2);
      y0 += t0;
      y1 += t1;
      y2 += t2;
    }
  }
  if constexpr (DIAG_LAST) {
    double d0, d1, d2;
    blk6(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], REC ? rec_vec(rec, p) : x[p], d0,
         d1, d2);
    y0 = d0 + y0;
    y1 = d1 + y1;
    y2 = d2 + y2;
  }
}

}  // namespace pd
