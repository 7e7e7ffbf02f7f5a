// pd_step.hip -- per-step kernels of the implicit peridynamics path:
// fused force / Jacobian / breaking / fill, block-sliced SpMV with the CG
// recurrences and their deterministic reductions, position update, contact.
//
// Floating point follows the reference's host semantics bit for bit: the
// file is compiled with -ffp-contract=off (no FMA), IEEE division and
// square root, and the two float roundings of cg_solver.cuh:256,264.
#include <math.h>

#include "pd_bond.h"

namespace pd {

// ---------------------------------------------------------------------------
// Fused assembly: core/cg_solver.cuh:200-339 (calculatePDForceDerivative)
// + core/implicit_euler.cuh:19-86 / core/newmark_beta.cuh:19-86 (fill A, b)
// + the block-Jacobi inverse of A_pp.  One lane per particle; a wave is one
// 64-row slice, so every slot read / write is 64 consecutive words.  With
// shared reciprocals a lane first takes its list three slots per trip
// without swap-remove, up to the first trip holding a breaking bond; the
// sequential pass (the next bond's column and neighbour data fetched while
// the current bond is computed) does the rest.  FP64 bond blocks are not
// written here: the solve's first product recomputes and writes them
// (writing them here cost more than that recomputation, the stores adding
// to the arithmetic rather than overlapping it), or every matrix-free
// product recomputes them.
// ---------------------------------------------------------------------------
// Neighbour q's position, stretch limit and (fuse) start value from the
// assembly record.
template <bool FUSE>
__device__ __forceinline__ void gather_asm(const AsmArgs& a, int q, double4& xq, double& sq,
                                           double4& us) {
  const dbl2v* r = (const dbl2v*)(a.arec + 8 * (int64_t)q);
  const dbl2v r0 = r[0], r1 = r[1];
  xq = make_double4(r0.x, r0.y, r1.x, 0.0);
  sq = r1.y;
  if constexpr (FUSE) {
    const dbl2v r2 = r[2], r3 = r[3];
    us = make_double4(r2.x, r2.y, r3.x, 0.0);
  }
}

// MIXED: the block-Jacobi inverse is written rounded to float, the
// diagonal in both precisions, and the start is the mixed solve's; the
// float bond blocks are left to the solve's first inner product.
// Neighbour data come from the step's assembly record (position, stretch
// limit and the solver's start value side by side, k_asm_prep).  SHARED:
// the quotients by the float-rounded length (bond_dir), by the reference
// length (dij / d0, c / d0), by the current length (d0 / dij) and by the
// radius (the volume scale) share reciprocals and the square root is
// sqrt_nr; both variants are launched and the one that does not match bit 0
// of the range flag (k_asm_prep) returns at once.  The SHARED variant sets
// bit 1 when a live bond has zero length (the matrix-free products then
// leave the shared forms).
// FUSE and GAP: a.fuse and a.track_gap as compile-time choices.
template <bool MIXED, bool SHARED, bool FUSE, bool GAP>
__global__ __launch_bounds__(PD_BLOCK) void k_assemble(AsmArgs a) {
  if (((a.counters->range_bad & 1u) == 0) != SHARED) return;
  if (a.consts) {
    a.dt = a.consts->dt;
    a.dt2 = a.consts->dt2;
    a.clock = a.consts->clock;
    a.gravity = a.consts->gravity;
  }
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile * PD_BLOCK + threadIdx.x;
  unsigned long long broken = 0, live = 0;
  bool zero_len = false;  // SHARED: a bond of zero current length
  double t_rr = 0.0, t_rz = 0.0, t_r32 = 0.0, t_bb = 0.0;  // fused residual partials
  if (p < a.n && a.act[p] != PD_INVISIBLE) {
    const pd_mat_dev mt = a.mat[a.body[p]];
    const int s = p >> 6;
    const int lane = p & 63;
    const int64_t base = a.slice_off[s];
    const double4 pos_p = a.x[p];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    int top = a.top[p];
    const int otop = top;
    const double stretch = a.s[p];
    const float stretch_f = (float)stretch;
    double smin = 0.0;
    double ssum = 0.0, smax = -INFINITY;
    double rb0 = 0.0, rb1 = 0.0, rb2 = 0.0;  // fused residual: bond terms
    double k00 = 0.0, k01 = 0.0, k02 = 0.0, k11 = 0.0, k12 = 0.0, k22 = 0.0;
    double gmin = INFINITY;
    int gq = -1;
    const double hr = mt.horizon - mt.radius;
    const double yr = SHARED ? rcp_nr(mt.radius) : 0.0;  // 1 / radius for vscale
    // A bond's geometry, stretch and breaking threshold, cg_solver.cuh:253-264
    // (yd: SHARED, 1 / d0 for dij / d0 and c / d0); the bond breaks when
    // drdivd > thr.
    auto measure = [&](const double4& xq, double sqv, double d0, double& e0, double& e1,
                       double& e2, double& dij, double& yd, double& drdivd, double& thr) {
      bond_dir<SHARED>(pos_p, xq, e0, e1, e2, dij);
      if constexpr (SHARED) zero_len = zero_len || dij == 0.0;
      yd = 0.0;
      if constexpr (SHARED) yd = rcp_nr(d0);
      drdivd = (SHARED ? div_rcp(dij, d0, yd) : dij / d0) - 1.0;
      // cg_solver.cuh:264: threshold (float)max(s_p, s_q)
      const float sq_f = (float)sqv;
      thr = (double)(stretch_f > sq_f ? stretch_f : sq_f);
    };
    // GAP: the bond nearest its threshold so far (the first one on ties)
    auto track_gap = [&](int q, double drdivd, double thr) {
      if constexpr (GAP) {
        const double gap = fabs(drdivd - thr);
        if (gap < gmin) {
          gmin = gap;
          gq = q;
        }
      }
    };
    // A live bond's force, Jacobian, diagonal and residual terms,
    // cg_solver.cuh:277-319.
    auto accumulate = [&](const double4& u, double d0, double e0, double e1, double e2,
                          double dij, double yd, double drdivd) {
      double vscale = 1.0;
      if (d0 > hr)
        vscale = 0.5 + (SHARED ? div_rcp((mt.horizon - d0) / 2.0, mt.radius, yr)
                               : (mt.horizon - d0) / 2.0 / mt.radius);
      if (drdivd < smin) smin = drdivd;
      ssum += drdivd;
      smax = drdivd > smax ? drdivd : smax;
      const double kk = drdivd * vscale;
      f0 += kk * e0;
      f1 += kk * e1;
      f2 += kk * e2;
      double F00, F01, F02, F11, F12, F22;
      const double sc = SHARED ? div_rcp(mt.c_bond, d0, yd) : mt.c_bond / d0;
      bond_jac_sc<SHARED>(e0, e1, e2, dij, d0, sc, F00, F01, F02, F11, F12, F22);
      k00 += F00;
      k01 += F01;
      k02 += F02;
      k11 += F11;
      k12 += F12;
      k22 += F22;
      const double b00 = (F00 * -1.0) * a.dt2, b01 = (F01 * -1.0) * a.dt2,
                   b02 = (F02 * -1.0) * a.dt2, b11 = (F11 * -1.0) * a.dt2,
                   b12 = (F12 * -1.0) * a.dt2, b22 = (F22 * -1.0) * a.dt2;
      // The blocks are written by the solve's first product (FP64:
      // k_cg_spmv_store or k_cg_init's storing variant; mixed, rounded to
      // float: k_mx_spmv_store) or, matrix-free, recomputed by every product;
      // here they feed the first residual of the FP64 system
      if constexpr (FUSE) {
        rb0 += b00 * u.x + b01 * u.y + b02 * u.z;
        rb1 += b01 * u.x + b11 * u.y + b12 * u.z;
        rb2 += b02 * u.x + b12 * u.y + b22 * u.z;
      }
    };
    int bnd = 0;
    // With shared reciprocals: first an optimistic pass over the list as
    // it stands, OPT slots per trip (independent chains for the
    // scheduler) and no swap-remove.  Breaking bonds are rare (tens per
    // step among 10^8): a lane stops at the first trip that holds one and
    // the sequential pass below takes over at that trip's first slot, with
    // the sums of the slots before it -- those of the sequential pass, which
    // reaches that slot with the same bonds in the same order.
    constexpr int OPT = SHARED ? 3 : 0;
    if constexpr (OPT > 0) {
      bool brk = false;
      for (; bnd + OPT - 1 <= top; bnd += OPT) {
        int qv[OPT];
        double lv[OPT], sv[OPT];
        double4 xv[OPT], uv[OPT];
#pragma unroll
        for (int k = 0; k < OPT; ++k) {
          qv[k] = ntl(a.col + (base + bnd + k) * 64 + lane);
          lv[k] = ntl(a.d0 + (base + bnd + k) * 64 + lane);
        }
#pragma unroll
        for (int k = 0; k < OPT; ++k) gather_asm<FUSE>(a, qv[k], xv[k], sv[k], uv[k]);
        double e0[OPT], e1[OPT], e2[OPT], dij[OPT], yd[OPT], dr[OPT], th[OPT];
#pragma unroll
        for (int k = 0; k < OPT; ++k) {
          measure(xv[k], sv[k], lv[k], e0[k], e1[k], e2[k], dij[k], yd[k], dr[k], th[k]);
          brk = brk || dr[k] > th[k];
        }
        if (brk) break;
#pragma unroll
        for (int k = 0; k < OPT; ++k) {
          track_gap(qv[k], dr[k], th[k]);
          accumulate(uv[k], lv[k], e0[k], e1[k], e2[k], dij[k], yd[k], dr[k]);
        }
      }
      for (; !brk && bnd <= top; ++bnd) {
        const int q = ntl(a.col + (base + bnd) * 64 + lane);
        const double d0 = ntl(a.d0 + (base + bnd) * 64 + lane);
        double4 xq, u;
        double sqv;
        gather_asm<FUSE>(a, q, xq, sqv, u);
        double e0, e1, e2, dij, yd, drdivd, thr;
        measure(xq, sqv, d0, e0, e1, e2, dij, yd, drdivd, thr);
        brk = drdivd > thr;
        if (brk) break;
        track_gap(q, drdivd, thr);
        accumulate(u, d0, e0, e1, e2, dij, yd, drdivd);
      }
    }
    // Sequential pass with swap-remove, in software pipeline: the neighbour
    // data of slot bnd + 1 and the column of slot bnd + 2 are in flight while
    // slot bnd is computed, so neither the column load nor the gathers it
    // feeds wait in line
    int q_n = 0, q_nn = 0;
    double4 xq_n = make_double4(0, 0, 0, 0);
    double4 us_n = make_double4(0, 0, 0, 0);  // fused residual: xs of the neighbour
    double sq_n = 0.0;
    double l_n = 0.0;  // the slot's reference length
    if (bnd <= top) {
      q_n = a.col[(base + bnd) * 64 + lane];
      l_n = a.d0[(base + bnd) * 64 + lane];
      gather_asm<FUSE>(a, q_n, xq_n, sq_n, us_n);
    }
    if (bnd + 1 <= top) q_nn = a.col[(base + bnd + 1) * 64 + lane];
    while (bnd <= top) {
      const int64_t slot = base + bnd;
      const int q = q_n;
      const double4 xq = xq_n;
      const double d0 = l_n;
      const double sqv = sq_n;
      const double4 u = us_n;
      if (bnd + 1 <= top) {
        q_n = q_nn;
        l_n = a.d0[(slot + 1) * 64 + lane];
        gather_asm<FUSE>(a, q_n, xq_n, sq_n, us_n);
      }
      if (bnd + 2 <= top) q_nn = a.col[(slot + 2) * 64 + lane];
      double e0, e1, e2, dij, yd, drdivd, thr;
      measure(xq, sqv, d0, e0, e1, e2, dij, yd, drdivd, thr);
      track_gap(q, drdivd, thr);
      if (drdivd > thr) {
        // swap-remove: the last live bond takes this slot, walker stays
        const int64_t last_slot = (base + top) * 64 + lane;
        const int last = a.col[last_slot];
        a.col[slot * 64 + lane] = last;
        a.col[last_slot] = q;
        {  // the reference lengths move with their columns
          const double dl = a.d0[last_slot];
          a.d0[last_slot] = a.d0[slot * 64 + lane];
          a.d0[slot * 64 + lane] = dl;
        }
        --top;
        // slot bnd again, now holding `last`: restart the pipeline there
        q_n = last;
        l_n = a.d0[slot * 64 + lane];
        gather_asm<FUSE>(a, last, xq_n, sq_n, us_n);
        if (bnd + 1 <= top) q_nn = a.col[(slot + 1) * 64 + lane];
        continue;
      }
      accumulate(u, d0, e0, e1, e2, dij, yd, drdivd);
      ++bnd;
    }
    // A_pp = K_pp dt^2 + m I
    const double d00 = k00 * a.dt2 + mt.mass;
    const double d01 = k01 * a.dt2 + 0.0;
    const double d02 = k02 * a.dt2 + 0.0;
    const double d11 = k11 * a.dt2 + mt.mass;
    const double d12 = k12 * a.dt2 + 0.0;
    const double d22 = k22 * a.dt2 + mt.mass;
    double* dgo = a.diag + (int64_t)s * 6 * 64 + lane;
    dgo[0 * 64] = d00;
    dgo[1 * 64] = d01;
    dgo[2 * 64] = d02;
    dgo[3 * 64] = d11;
    dgo[4 * 64] = d12;
    dgo[5 * 64] = d22;
    if constexpr (MIXED) {
      float* dgf = a.diag32 + (int64_t)s * 6 * 64 + lane;
      dgf[0 * 64] = (float)d00;
      dgf[1 * 64] = (float)d01;
      dgf[2 * 64] = (float)d02;
      dgf[3 * 64] = (float)d11;
      dgf[4 * 64] = (float)d12;
      dgf[5 * 64] = (float)d22;
    }
    // block-Jacobi inverse by cofactors (six distinct entries)
    float m32[6];
    double m64[6];
    {
      const double c00 = d11 * d22 - d12 * d12;
      const double c01 = d02 * d12 - d01 * d22;
      const double c02 = d01 * d12 - d02 * d11;
      const double c11 = d00 * d22 - d02 * d02;
      const double c12 = d01 * d02 - d00 * d12;
      const double c22 = d00 * d11 - d01 * d01;
      const double det = (d00 * c00 + d01 * c01) + d02 * c02;
      const double id = 1.0 / det;
      if constexpr (MIXED) {
        m32[0] = (float)(c00 * id);
        m32[1] = (float)(c01 * id);
        m32[2] = (float)(c02 * id);
        m32[3] = (float)(c11 * id);
        m32[4] = (float)(c12 * id);
        m32[5] = (float)(c22 * id);
        float* mo = a.minv32 + (int64_t)s * 6 * 64 + lane;
#pragma unroll
        for (int c = 0; c < 6; ++c) mo[c * 64] = m32[c];
      } else {
        m64[0] = c00 * id;
        m64[1] = c01 * id;
        m64[2] = c02 * id;
        m64[3] = c11 * id;
        m64[4] = c12 * id;
        m64[5] = c22 * id;
        double* mo = a.minv + (int64_t)s * 6 * 64 + lane;
#pragma unroll
        for (int c = 0; c < 6; ++c) mo[c * 64] = m64[c];
      }
    }
    f0 *= mt.c_bond;
    f1 *= mt.c_bond;
    f2 *= mt.c_bond;
    const double ss0 = (10.0 - 9.0 * a.clock) * a.s00[p];
    a.s_new[p] = ss0 - 0.25 * smin;
    f1 -= (a.clock * a.gravity * mt.mass);
    if (a.contact) {
      const double4 c = a.fc[p];
      const double len = sqrt(c.x * c.x + c.y * c.y + c.z * c.z);
      if (len > 1e-10) {
        f0 += c.x;
        f1 += c.y;
        f2 += c.z;
      }
    }
    a.f[p] = make_double4(f0, f1, f2, 0.0);
    const double4 vp = a.v[p];
    double4 bo;
    if (a.integrator == 1) {
      const double hdt = 0.5 * a.dt;
      bo.x = a.dt * (mt.mass * vp.x + hdt * f0);
      bo.y = a.dt * (mt.mass * vp.y + hdt * f1);
      bo.z = a.dt * (mt.mass * vp.z + hdt * f2);
    } else {
      bo.x = mt.mass * vp.x + a.dt * f0;
      bo.y = mt.mass * vp.y + a.dt * f1;
      bo.z = mt.mass * vp.z + a.dt * f2;
    }
    bo.w = 0.0;
    a.b[p] = bo;
    if constexpr (FUSE) {
      // y = D x_p + (bond terms); r = S(b - y); the solver's start
      const double4 xp = a.xs[p];
      const double y0 = (d00 * xp.x + d01 * xp.y + d02 * xp.z) + rb0;
      const double y1 = (d01 * xp.x + d11 * xp.y + d12 * xp.z) + rb1;
      const double y2 = (d02 * xp.x + d12 * xp.y + d22 * xp.z) + rb2;
      const bool fr = a.bc[p] == 0;
      const double r0 = fr ? bo.x - y0 : 0.0;
      const double r1 = fr ? bo.y - y1 : 0.0;
      const double r2 = fr ? bo.z - y2 : 0.0;
      const double sb0 = fr ? bo.x : 0.0, sb1 = fr ? bo.y : 0.0, sb2 = fr ? bo.z : 0.0;
      t_rr = (r0 * r0 + r1 * r1) + r2 * r2;
      t_bb = (sb0 * sb0 + sb1 * sb1) + sb2 * sb2;
      if constexpr (MIXED) {
        const float h0 = (float)r0, h1 = (float)r1, h2 = (float)r2;
        float z0 = h0, z1 = h1, z2 = h2;
        if (a.pc) {  // z = M32 r32, the float block-Jacobi inverse just written
          z0 = m32[0] * h0 + m32[1] * h1 + m32[2] * h2;
          z1 = m32[1] * h0 + m32[3] * h1 + m32[4] * h2;
          z2 = m32[2] * h0 + m32[4] * h1 + m32[5] * h2;
        }
        a.r32[p] = make_float4(h0, h1, h2, 0.0f);
        a.p32[p] = make_float4(z0, z1, z2, 0.0f);
        a.d32[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        t_rz = ((double)h0 * (double)z0 + (double)h1 * (double)z1) + (double)h2 * (double)z2;
        t_r32 = ((double)h0 * (double)h0 + (double)h1 * (double)h1) + (double)h2 * (double)h2;
      } else {
        double z0 = r0, z1 = r1, z2 = r2;
        if (a.pc) {  // z = M r, the block-Jacobi inverse just written
          z0 = m64[0] * r0 + m64[1] * r1 + m64[2] * r2;
          z1 = m64[1] * r0 + m64[3] * r1 + m64[4] * r2;
          z2 = m64[2] * r0 + m64[4] * r1 + m64[5] * r2;
        }
        a.r64[p] = make_double4(r0, r1, r2, 0.0);
        if (a.p64) a.p64[p] = make_double4(z0, z1, z2, 0.0);  // null: matrix-free
        if (a.rec) {
          double* rp = a.rec + 6 * (int64_t)p;
          rp[3] = z0;
          rp[4] = z1;
          rp[5] = z2;
        }
        t_rz = (r0 * z0 + r1 * z1) + r2 * z2;
      }
    }
    if (!FUSE && a.bc[p]) {  // fused: k_asm_prep wrote it
      const double4 vb = a.vbc[p];
      if (a.integrator == 1)
        a.sol[p] = make_double4(a.dt * vb.x, a.dt * vb.y, a.dt * vb.z, 0.0);
      else
        a.sol[p] = make_double4(vb.x, vb.y, vb.z, 0.0);
    }
    if (top != otop) {
      a.top[p] = top;
      broken = (unsigned long long)(otop - top);
    }
    live = (unsigned long long)(top + 1);
    if (a.bond_stats)
      a.bond_stats[p] = make_float2(top >= 0 ? (float)(ssum / (top + 1)) : 0.0f,
                                    top >= 0 ? (float)smax : 0.0f);
    if constexpr (GAP) {
      a.gap_val[p] = gmin;
      a.gap_q[p] = gq;
    }
  } else if (p < a.n) {
    if constexpr (GAP) {
      a.gap_val[p] = INFINITY;
      a.gap_q[p] = -1;
    }
    if constexpr (FUSE) {
      if constexpr (MIXED) {
        const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        a.r32[p] = z4;
        a.p32[p] = z4;
        a.d32[p] = z4;
      } else {
        const double4 z4 = make_double4(0.0, 0.0, 0.0, 0.0);
        a.r64[p] = z4;
        if (a.p64) a.p64[p] = z4;
        if (a.rec) {
          double* rp = a.rec + 6 * (int64_t)p;
          rp[3] = 0.0;
          rp[4] = 0.0;
          rp[5] = 0.0;
        }
      }
    }
  }
  if constexpr (FUSE) {  // the start kernel's tile partials, all threads
    __shared__ double lds4[4];
    const int ntiles = gridDim.x;
    if constexpr (MIXED) {
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
    } else {
      const double v0 = tile_sum(t_rz, lds4);
      const double v1 = tile_sum(t_rr, lds4);
      const double v2 = tile_sum(t_bb, lds4);
      if (threadIdx.x == 0) {
        a.part[tile] = v0;
        a.part[ntiles + tile] = v1;
        a.part[2 * (int64_t)ntiles + tile] = v2;
      }
    }
  }
  // integer counters: order-free
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    broken += __shfl_xor(broken, m, 64);
    live += __shfl_xor(live, m, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    if (broken) atomicAdd(&a.counters->broken, broken);
    if (live) atomicAdd(&a.counters->live, live);
  }
  if constexpr (SHARED)
    if (__any(zero_len) && (threadIdx.x & 63) == 0) atomicOr(&a.counters->range_bad, 2u);
}

template <bool MIXED, bool FUSE, bool GAP>
void launch_assemble_t(const AsmArgs& a, int nb, hipStream_t s) {
  // both variants; the one the range flag does not select returns at once
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_assemble<MIXED, true, FUSE, GAP>), dim3(nb), dim3(PD_BLOCK),
                     0, s, a);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_assemble<MIXED, false, FUSE, GAP>), dim3(nb), dim3(PD_BLOCK),
                     0, s, a);
}

void launch_assemble(const AsmArgs& a, hipStream_t s) {
  const int nb = (a.n + PD_BLOCK - 1) / PD_BLOCK;
  if (!nb) return;
  const int k = (a.mixed ? 4 : 0) + (a.fuse ? 2 : 0) + (a.track_gap ? 1 : 0);
  switch (k) {
    case 0: launch_assemble_t<false, false, false>(a, nb, s); break;
    case 1: launch_assemble_t<false, false, true>(a, nb, s); break;
    case 2: launch_assemble_t<false, true, false>(a, nb, s); break;
    case 3: launch_assemble_t<false, true, true>(a, nb, s); break;
    case 4: launch_assemble_t<true, false, false>(a, nb, s); break;
    case 5: launch_assemble_t<true, false, true>(a, nb, s); break;
    case 6: launch_assemble_t<true, true, false>(a, nb, s); break;
    default: launch_assemble_t<true, true, true>(a, nb, s); break;
  }
}

// The step's assembly record, ahead of the assembly: per particle
// (x, y, z, s, xs.x, xs.y, xs.z, 0) -- position, stretch limit and the
// solver's start value, one 64-byte gather per bond in the assembly.  With
// fuse, the Dirichlet rows' start values are set first (vbc, or dt vbc for
// Newmark-beta).  The position part of the matrix-free record is written
// here too, and bit 0 of the range flag of the shared-reciprocal quotients:
// set when a particle's coordinate, a reference length (d0_bad, from
// k_bond_d0) or a body's bond stiffness, horizon or radius (c_bad) lies
// outside the ranges of range_ok_*.
__global__ __launch_bounds__(PD_BLOCK) void k_asm_prep(AsmPrepArgs a) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  bool bad = false;
  if (p < a.n) {
    const double4 xp = a.x[p];
    double4 xs = a.sol[p];
    const bool vis = a.act[p] != PD_INVISIBLE;
    if (a.fuse && a.bc[p] && vis) {
      const double dt = a.consts ? a.consts->dt : a.dt;
      const double4 vb = a.vbc[p];
      xs = a.integrator == 1 ? make_double4(dt * vb.x, dt * vb.y, dt * vb.z, 0.0)
                             : make_double4(vb.x, vb.y, vb.z, 0.0);
      a.sol[p] = xs;
    }
    dbl2v* r = (dbl2v*)(a.arec + 8 * (int64_t)p);
    r[0] = dbl2v{xp.x, xp.y};
    r[1] = dbl2v{xp.z, a.s[p]};
    r[2] = dbl2v{xs.x, xs.y};
    r[3] = dbl2v{xs.z, 0.0};
    if (a.rec) {
      double* rp = a.rec + 6 * (int64_t)p;
      rp[0] = xp.x;
      rp[1] = xp.y;
      rp[2] = xp.z;
    }
    // every particle: an invisible one may still be the neighbour of a
    // visible one whose bonds were built before its activity changed
    bad = !(range_ok_coord(xp.x) && range_ok_coord(xp.y) && range_ok_coord(xp.z));
  }
  if (p == 0 && (a.c_bad || *a.d0_bad)) bad = true;
  if (__any(bad) && (threadIdx.x & 63) == 0) atomicOr(&a.counters->range_bad, 1u);
}

void launch_asm_prep(const AsmPrepArgs& a, hipStream_t s) {
  const int nb = (a.n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_asm_prep, dim3(nb), dim3(PD_BLOCK), 0, s, a);
}

__global__ void k_reset_counters(pd_step_counters* c) {
  if (threadIdx.x == 0) {
    c->broken = 0;
    c->live = 0;
    c->range_bad = 0;
  }
}

void launch_reset_counters(pd_step_counters* c, hipStream_t s) {
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, c);
}

// Deterministic arg-min of the per-row gaps: smallest gap, then smallest p.
__global__ __launch_bounds__(PD_BLOCK) void k_gap_partial(const double* gv,
                                                         const int32_t* gq, int n,
                                                         double* pv, int32_t* pp,
                                                         int32_t* pq) {
  __shared__ double sv[PD_BLOCK];
  __shared__ int sp[PD_BLOCK];
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  sv[threadIdx.x] = i < n ? gv[i] : INFINITY;
  sp[threadIdx.x] = i < n ? i : 0x7fffffff;
  __syncthreads();
  for (int w = PD_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double a = sv[threadIdx.x], b = sv[threadIdx.x + w];
      const int ia = sp[threadIdx.x], ib = sp[threadIdx.x + w];
      if (b < a || (b == a && ib < ia)) {
        sv[threadIdx.x] = b;
        sp[threadIdx.x] = ib;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pp[blockIdx.x] = sp[0];
    pq[blockIdx.x] = (sp[0] < n) ? gq[sp[0]] : -1;
  }
}

__global__ __launch_bounds__(PD_BLOCK) void k_gap_final(const double* pv,
                                                       const int32_t* pp,
                                                       const int32_t* pq, int g,
                                                       double* out_val,
                                                       int32_t* out_pq) {
  __shared__ double sv[PD_BLOCK];
  __shared__ int sp[PD_BLOCK];
  __shared__ int sq[PD_BLOCK];
  double bv = INFINITY;
  int bp = 0x7fffffff, bq = -1;
  for (int k = threadIdx.x; k < g; k += PD_BLOCK) {
    if (pv[k] < bv || (pv[k] == bv && pp[k] < bp)) {
      bv = pv[k];
      bp = pp[k];
      bq = pq[k];
    }
  }
  sv[threadIdx.x] = bv;
  sp[threadIdx.x] = bp;
  sq[threadIdx.x] = bq;
  __syncthreads();
  for (int w = PD_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double a = sv[threadIdx.x], b = sv[threadIdx.x + w];
      const int ia = sp[threadIdx.x], ib = sp[threadIdx.x + w];
      if (b < a || (b == a && ib < ia)) {
        sv[threadIdx.x] = b;
        sp[threadIdx.x] = ib;
        sq[threadIdx.x] = sq[threadIdx.x + w];
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out_val = sv[0];
    out_pq[0] = sp[0] == 0x7fffffff ? -1 : sp[0];
    out_pq[1] = sq[0];
  }
}

void launch_gap_reduce(const double* gap_val, const int32_t* gap_q, int32_t n,
                       double* part_val, int32_t* part_p, int32_t* part_q,
                       double* out_val, int32_t* out_pq, hipStream_t s) {
  const int g = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (g == 0) return;
  hipLaunchKernelGGL(k_gap_partial, dim3(g), dim3(PD_BLOCK), 0, s, gap_val, gap_q, n,
                     part_val, part_p, part_q);
  hipLaunchKernelGGL(k_gap_final, dim3(1), dim3(PD_BLOCK), 0, s, part_val, part_p, part_q,
                     g, out_val, out_pq);
}

// ---------------------------------------------------------------------------
// SpMV row: core/cg_solver.cuh:343-383 computeAx.
//   y_p = A_pp x_p, then y_p += dot(rows of A_{p,b}, x_{q_b}) for the live
//   slots in list order; every block product is (a0 x0 + a1 x1) + a2 x2.
// ---------------------------------------------------------------------------
// U slots per trip: all their loads are issued before any product (the
// summation order does not depend on U).  DIAG_LAST: the bond terms are
// summed from 0.0 and the diagonal term is added last (the solver's first
// residual, the order the assembly pass forms it in).
template <int U, bool DIAG_LAST = false>
__device__ __forceinline__ void spmv_row(const SpmvView& A, const double4* __restrict__ x,
                                         int p, double& y0, double& y1, double& y2) {
  const int s = p >> 6;
  const int lane = p & 63;
  const double* dg = A.diag + (int64_t)s * 6 * 64 + lane;
  if constexpr (DIAG_LAST) {
    y0 = 0.0;
    y1 = 0.0;
    y2 = 0.0;
  } else {
    blk6(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], x[p], y0, y1, y2);
  }
  const int64_t base = A.slice_off[s];
  const int top = A.top[p];
  const int* __restrict__ cp = A.col + base * 64 + lane;
  // three planes of value pairs per slot (16 bytes per lane and load)
  const vpair* __restrict__ vp = (const vpair*)(A.val + base * 6 * 64) + lane;
  int bnd = 0;
  for (; bnd + U - 1 <= top; bnd += U) {
    int q[U];
    double v[U][6];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = ntl(cp + (int64_t)(bnd + u) * 64);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const vpair* vv = vp + (int64_t)(bnd + u) * 192;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const vpair t = __builtin_nontemporal_load(vv + c * 64);
        v[u][2 * c] = t.x;
        v[u][2 * c + 1] = t.y;
      }
    }
    double4 xx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) xx[u] = x[q[u]];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double t0, t1, t2;
      blk6(v[u][0], v[u][1], v[u][2], v[u][3], v[u][4], v[u][5], xx[u], t0, t1, t2);
      y0 += t0;
      y1 += t1;
      y2 += t2;
    }
  }
  for (; bnd <= top; ++bnd) {
    const int q0 = ntl(cp + (int64_t)bnd * 64);
    const vpair* v0 = vp + (int64_t)bnd * 192;
    const vpair a0 = __builtin_nontemporal_load(v0), a1 = __builtin_nontemporal_load(v0 + 64),
                a2 = __builtin_nontemporal_load(v0 + 128);
    double t0, t1, t2;
    blk6(a0.x, a0.y, a1.x, a1.y, a2.x, a2.y, x[q0], t0, t1, t2);
    y0 += t0;
    y1 += t1;
    y2 += t2;
  }
  if constexpr (DIAG_LAST) {
    double d0, d1, d2;
    blk6(dg[0], dg[64], dg[128], dg[192], dg[256], dg[320], x[p], d0, d1, d2);
    y0 = d0 + y0;
    y1 = d1 + y1;
    y2 = d2 + y2;
  }
}

// U > 0: every block stored, U slots per trip; U = 0: every block
// recomputed, three slots per trip, with the gathers from the record when
// REC and, when SHARED_OK and the assembly's range flag allows (a
// wave-uniform choice), shared reciprocals.  The CG SpMV takes SHARED_OK;
// in the plain SpMV and the solver start a kernel holding both quotient
// variants ran 20 % slower than either alone, so those divide one quotient
// at a time.
template <int U, bool DIAG_LAST = false, bool REC = false, bool SHARED_OK = false>
__device__ __forceinline__ void row_of(const SpmvView& A, const double4* __restrict__ x, int p,
                                       double& y0, double& y1, double& y2) {
  if constexpr (U == 0) {
    if (SHARED_OK && A.range_bad && *A.range_bad == 0)
      spmv_row_mf<3, DIAG_LAST, REC, true>(A, x, p, y0, y1, y2, A.rec);
    else
      spmv_row_mf<3, DIAG_LAST, REC, false>(A, x, p, y0, y1, y2, A.rec);
  } else {
    spmv_row<U, DIAG_LAST>(A, x, p, y0, y1, y2);
  }
}

template <int U>
__global__ __launch_bounds__(PD_BLOCK) void k_spmv_plain(SpmvView A,
                                                        const double4* __restrict__ x,
                                                        double4* __restrict__ y) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= A.n) return;
  double y0, y1, y2;
  row_of<U>(A, x, p, y0, y1, y2);
  y[p] = make_double4(y0, y1, y2, 0.0);
}

// Self-check of the shared-reciprocal quotients against IEEE division, and
// of sqrt_nr against IEEE sqrt, over random operands spanning the ranges
// the range flag admits (odd samples: a coordinate difference over a float
// bond length; 2 mod 4: a current bond length over a reference length;
// 0 mod 16: a bond stiffness over a reference length; 8 mod 16: half of a
// horizon less a reference length over a radius; 4 mod 16: a reference
// length over a current length; 12 mod 16: the square root of a squared
// bond length); the exponent ends are drawn often.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
__device__ __forceinline__ double draw_f64(uint64_t h, int elo, int ehi) {
  const int span = ehi - elo + 1;
  const int pick = (int)((h >> 52) & 7);
  const int e = pick == 0 ? elo : pick == 1 ? ehi : elo + (int)((h >> 12) % (uint64_t)span);
  const uint64_t bits = ((h >> 63) << 63) | ((uint64_t)(e + 1023) << 52) |
                        (mix64(h) & 0xfffffffffffffull);
  return __longlong_as_double((long long)bits);
}
__global__ __launch_bounds__(PD_BLOCK) void k_check_division(int64_t samples, uint64_t seed,
                                                             unsigned long long* bad) {
  const int64_t i = (int64_t)blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i >= samples) return;
  const uint64_t h1 = mix64(seed ^ (2 * (uint64_t)i)), h2 = mix64(seed ^ (2 * (uint64_t)i + 1));
  double a, b;
  if (i & 1) {
    a = (h1 & 15) == 0 ? 0.0 : draw_f64(h1, -352, 100);
    const float small_f = (float)1.0e-10;
    float bf = (h2 & 15) == 0 ? small_f : fabsf((float)draw_f64(h2, -34, 101));
    if (!(bf >= small_f)) bf = small_f;
    b = (double)bf;
  } else if ((i & 3) == 2) {  // a current bond length over a reference length
    a = (h1 & 15) == 0 ? 0.0 : fabs(draw_f64(h1, -537, 101));
    b = fabs(draw_f64(h2, -400, 99));
  } else if ((i & 15) == 0) {  // a bond stiffness over a reference length
    a = fabs(draw_f64(h1, -300, 299));
    b = fabs(draw_f64(h2, -400, 99));
  } else if ((i & 15) == 8) {  // (horizon - d0) / 2 over a radius
    a = (h1 & 15) == 0 ? 0.0 : draw_f64(h1, -453, 299);
    b = fabs(draw_f64(h2, -300, 299));
  } else if ((i & 15) == 4) {  // a reference length over a current length
    a = fabs(draw_f64(h1, -400, 99));
    b = fabs(draw_f64(h2, -352, 101));
  } else {  // the square root of a squared bond length
    const double x = fabs(draw_f64(h1, -704, 203));
    if (__double_as_longlong(sqrt(x)) != __double_as_longlong(sqrt_nr(x))) atomicAdd(bad, 1ull);
    return;
  }
  const double ieee = a / b;
  const double shared = div_rcp(a, b, rcp_nr(b));
  if (__double_as_longlong(ieee) != __double_as_longlong(shared)) atomicAdd(bad, 1ull);
}

void launch_check_division(int64_t samples, uint64_t seed, unsigned long long* bad,
                           hipStream_t s) {
  const int64_t nb = (samples + PD_BLOCK - 1) / PD_BLOCK;
  if (nb > 0) hipLaunchKernelGGL(k_check_division, dim3((unsigned)nb), dim3(PD_BLOCK), 0, s,
                                 samples, seed, bad);
}

void launch_spmv_plain(const SpmvView& A, const double4* x, double4* y, hipStream_t s) {
  const int nb = (A.n + PD_BLOCK - 1) / PD_BLOCK;
  if (!nb) return;
  if (A.mf) {
    hipLaunchKernelGGL(k_spmv_plain<0>, dim3(nb), dim3(PD_BLOCK), 0, s, A, x, y);
    return;
  }
  switch (A.unroll) {
    case 2: hipLaunchKernelGGL(k_spmv_plain<2>, dim3(nb), dim3(PD_BLOCK), 0, s, A, x, y); break;
    case 8: hipLaunchKernelGGL(k_spmv_plain<8>, dim3(nb), dim3(PD_BLOCK), 0, s, A, x, y); break;
    default: hipLaunchKernelGGL(k_spmv_plain<4>, dim3(nb), dim3(PD_BLOCK), 0, s, A, x, y);
  }
}

// ---------------------------------------------------------------------------
// CG (CUSP krylov::cg structure, core/simulator.cu:763-770) with scalars
// kept on the device.  Every streaming kernel writes one partial sum per
// 256-row tile; the partials of a specimen are folded in the fixed order of
// the oracle's orc_dot, by a fold kernel with one workgroup per specimen
// (the start, beta plus the monitor test) or by every workgroup of the
// update for itself (alpha).  Kernel boundaries order the writes, so no
// fences are needed.  The kernels of a tile or specimen return at once when
// its solve is done.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void precond(const double* minv, int pc, int p, double r0,
                                        double r1, double r2, double& z0, double& z1,
                                        double& z2) {
  if (pc != 0) {  // block Jacobi (1) or the learned block (2)
    const double* m = minv + (int64_t)(p >> 6) * 6 * 64 + (p & 63);
    const double m00 = m[0], m01 = m[64], m02 = m[128], m11 = m[192], m12 = m[256],
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

__device__ __forceinline__ double term3(double a0, double a1, double a2, double b0,
                                        double b1, double b2) {
  return (a0 * b0 + a1 * b1) + a2 * b2;
}

__global__ void k_cg_setup(pd_cg_state* states, int nspec, double rtol, int max_iter) {
  for (int k = threadIdx.x; k < nspec; k += blockDim.x) {
    pd_cg_state* st = states + k;
    st->rz = 0.0;
    st->yp = 0.0;
    st->alpha = 0.0;
    st->beta = 0.0;
    st->rr = 0.0;
    st->bnorm = 0.0;
    st->tol = 0.0;
    st->rnorm = 0.0;
    st->rtol = rtol;
    st->iter = 0;
    st->done = 0;
    st->converged = 0;
    st->max_iter = max_iter;
    st->upd_it = -1;
  }
}

void launch_cg_setup(pd_cg_state* st, int32_t nspec, double rtol, int32_t max_iter,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_cg_setup, dim3(1), dim3(256), 0, s, st, nspec, rtol, max_iter);
}

// The CG direction's part of the matrix-free gather record (rec non-null)
__device__ __forceinline__ void rec_p(double* rec, int p, double z0, double z1, double z2) {
  if (rec) {
    double* rp = rec + 6 * (int64_t)p;
    rp[3] = z0;
    rp[4] = z1;
    rp[5] = z2;
  }
}

// y = A x0; r = S(b - y); z = M r; p = z; partials <r,z>, <r,r>, <Sb,Sb>.
// STORE: a stored-block FP64 solve whose assembly left the blocks to it:
// every block is recomputed from the positions, written to a.vals and used.
template <int U, bool STORE = false>
__global__ __launch_bounds__(PD_BLOCK) void k_cg_init(CgArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t_rz = 0.0, t_rr = 0.0, t_bb = 0.0;
  if (p < a.A.n) {
    double y0, y1, y2;
    if constexpr (STORE)
      spmv_row_mf<3, true, false, false, true>(a.A, a.x, p, y0, y1, y2, nullptr, a.vals);
    else
      row_of<U, true>(a.A, a.x, p, y0, y1, y2);
    const double4 bp = a.b[p];
    const bool fr = a.bc[p] == 0;
    const double r0 = fr ? bp.x - y0 : 0.0;
    const double r1 = fr ? bp.y - y1 : 0.0;
    const double r2 = fr ? bp.z - y2 : 0.0;
    const double sb0 = fr ? bp.x : 0.0, sb1 = fr ? bp.y : 0.0, sb2 = fr ? bp.z : 0.0;
    double z0, z1, z2;
    precond(a.minv, a.precond, p, r0, r1, r2, z0, z1, z2);
    a.r[p] = make_double4(r0, r1, r2, 0.0);
    a.p[p] = make_double4(z0, z1, z2, 0.0);
    rec_p(a.A.rec, p, z0, z1, z2);
    t_rz = term3(r0, r1, r2, z0, z1, z2);
    t_rr = term3(r0, r1, r2, r0, r1, r2);
    t_bb = term3(sb0, sb1, sb2, sb0, sb1, sb2);
  }
  const double v0 = tile_sum(t_rz, lds4);
  const double v1 = tile_sum(t_rr, lds4);
  const double v2 = tile_sum(t_bb, lds4);
  if (threadIdx.x == 0) {
    a.part[tile] = v0;
    a.part[ntiles + tile] = v1;
    a.part[2 * (int64_t)ntiles + tile] = v2;
  }
}

// one workgroup per specimen
// SETUP: the solver state is set up here (k_cg_setup's values, with rtol
// and max_iter), else k_cg_setup ran before and they are read from it.
template <bool SETUP>
__global__ __launch_bounds__(PD_BLOCK) void k_cg_init_fold(CgArgs a, int ntiles, double rtol,
                                                          int max_iter) {
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  const double rz = fold_partials(a.part + t0, nt, lds4);
  const double rr = fold_partials(a.part + ntiles + t0, nt, lds4);
  const double bb = fold_partials(a.part + 2 * (int64_t)ntiles + t0, nt, lds4);
  if (threadIdx.x == 0) {
    pd_cg_state* st = a.st + blockIdx.x;
    if constexpr (SETUP) {  // the state's start (k_cg_setup's)
      st->yp = 0.0;
      st->alpha = 0.0;
      st->beta = 0.0;
      st->rtol = rtol;
      st->done = 0;
      st->converged = 0;
      st->max_iter = max_iter;
      st->upd_it = -1;
    } else {
      rtol = st->rtol;
      max_iter = st->max_iter;
    }
    const double bnorm = sqrt(bb);
    const double tol = rtol * bnorm;
    const double rnorm = sqrt(rr);
    st->rz = rz;
    st->rr = rr;
    st->bnorm = bnorm;
    st->tol = tol;
    st->rnorm = rnorm;
    st->iter = 0;
    if (rnorm <= tol) {
      st->done = 1;
      st->converged = 1;
    } else if (max_iter <= 0) {
      st->done = 1;
    }
  }
}

// y = A p; partial <y,p>
template <int U>
__global__ __launch_bounds__(PD_BLOCK) void k_cg_spmv(CgArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (a.st[spec_of(a.batch, tile)].done) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t = 0.0;
  if (p < a.A.n) {
    double y0, y1, y2;
    if (U == 0 && a.A.rec)
      row_of<U, false, true, true>(a.A, a.p, p, y0, y1, y2);
    else
      row_of<U, false, false, true>(a.A, a.p, p, y0, y1, y2);
    a.y[p] = make_double4(y0, y1, y2, 0.0);
    const double4 pp = a.p[p];
    t = term3(y0, y1, y2, pp.x, pp.y, pp.z);
  }
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[tile] = v;
}

// The matrix-free CG product after an assembly: gathers from the record;
// with the range flag clear (the common case) shared reciprocals and
// sqrt_nr, four bonds per trip, else IEEE quotients, two bonds per trip
// (the general path fits in the common path's registers).
// The row loop waits on a dependent chain (column, then the neighbour's
// record) each trip, so the bonds per trip, not the waves per SIMD, hide
// its latency: at 1M particles 511 us per product with four bonds per trip
// (126 VGPRs, four waves per SIMD) against 562 with two (76 VGPRs, six
// waves), 592 with three, 569 with six and 557 with eight.
__global__ __launch_bounds__(PD_BLOCK) void k_cg_spmv_mf(CgArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  if (a.st[spec_of(a.batch, tile)].done) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t = 0.0;
  if (p < a.A.n) {
    double y0, y1, y2;
    if (*a.A.range_bad == 0)  // uniform over the grid
      spmv_row_mf<4, false, true, true>(a.A, a.p, p, y0, y1, y2, a.A.rec);
    else
      spmv_row_mf<2, false, true, false>(a.A, a.p, p, y0, y1, y2, a.A.rec);
    a.y[p] = make_double4(y0, y1, y2, 0.0);
    const double4 pp = rec_vec(a.A.rec, p);  // p lives in the record alone
    t = term3(y0, y1, y2, pp.x, pp.y, pp.z);
  }
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[tile] = v;
}

// The first CG product of a stored-block FP64 step whose assembly left the
// blocks out: every block is recomputed from the gather record (positions
// and p side by side) exactly as the assembly forms it, written to its slot
// and multiplied; the later products stream the written blocks.  Tiles of
// converged specimens write their blocks too but no product or partial.
__global__ __launch_bounds__(PD_BLOCK) void k_cg_spmv_store(CgArgs a) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const bool done = a.st[spec_of(a.batch, tile)].done;
  const int p = tile * PD_BLOCK + threadIdx.x;
  double t = 0.0;
  if (p < a.A.n) {
    double y0, y1, y2;
    if (a.A.range_bad && *a.A.range_bad == 0)
      spmv_row_mf<3, false, true, true, true>(a.A, a.p, p, y0, y1, y2, a.rec, a.vals);
    else
      spmv_row_mf<3, false, true, false, true>(a.A, a.p, p, y0, y1, y2, a.rec, a.vals);
    if (!done) {
      a.y[p] = make_double4(y0, y1, y2, 0.0);
      const double4 pp = a.p[p];
      t = term3(y0, y1, y2, pp.x, pp.y, pp.z);
    }
  }
  if (done) return;  // uniform over the tile
  const double v = tile_sum(t, lds4);
  if (threadIdx.x == 0) a.part[tile] = v;
}

// The blocks of the current bond lists from the positions, written to vals.
__global__ __launch_bounds__(PD_BLOCK) void k_write_blocks(SpmvView A, double* vals) {
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
    double e0, e1, e2, dij;
    bond_dir(pos_p, A.pos[A.col[i]], e0, e1, e2, dij);
    double F00, F01, F02, F11, F12, F22;
    bond_jac(e0, e1, e2, dij, A.d0[i], cb, F00, F01, F02, F11, F12, F22);
    store_block(vals, base + bnd, lane, (F00 * -1.0) * dt2, (F01 * -1.0) * dt2,
                (F02 * -1.0) * dt2, (F11 * -1.0) * dt2, (F12 * -1.0) * dt2, (F22 * -1.0) * dt2);
  }
}

// Iteration `it` of the CG loop: alpha = rz / <y,p>, formed by every
// workgroup from its specimen's <y,p> partials in the one fold order (so
// all hold the same value); then r -= alpha y (free rows); z = M r;
// partials <r,z>, <r,r> into the third and fourth partial rows (the first
// row is still being folded by other workgroups).  x += alpha p is left to
// k_cg_dir, which reads p anyway.  The specimen's first workgroup records
// alpha and that the update of iteration `it` ran.
__global__ __launch_bounds__(PD_BLOCK) void k_cg_update(CgArgs a, int it) {
  __shared__ double lds4[4];
  const int ntiles = gridDim.x;
  const int tile = xcd_tile(blockIdx.x, ntiles);
  const int s = spec_of(a.batch, tile);
  pd_cg_state* st = a.st + s;
  if (st->done) return;
  const int p = tile * PD_BLOCK + threadIdx.x;
  const bool row = p < a.A.n;
  // the row's loads go out ahead of the fold
  double4 rr = make_double4(0.0, 0.0, 0.0, 0.0), yy = rr;
  bool fr = false;
  if (row) {
    rr = a.r[p];
    fr = a.bc[p] == 0;
    if (fr) yy = a.y[p];
  }
  int t0, nt;
  spec_range_of(a.batch, s, ntiles, t0, nt);
  const double rz = st->rz;
  double alpha;
  if (nt > PD_ALPHA_FOLD_TILES) {
    alpha = st->alpha;  // k_cg_alpha folded it ahead of this launch
  } else {
    const double yp = fold_partials(a.part + t0, nt, lds4);
    alpha = rz / yp;
    if (tile == t0 && threadIdx.x == 0) {
      st->yp = yp;
      st->alpha = alpha;
    }
  }
  if (tile == t0 && threadIdx.x == 0) st->upd_it = it;
  double t_rz = 0.0, t_rr = 0.0;
  if (row) {
    if (fr) {
      const double na = -alpha;
      rr.x = na * yy.x + rr.x;
      rr.y = na * yy.y + rr.y;
      rr.z = na * yy.z + rr.z;
      a.r[p] = rr;
    }
    double z0, z1, z2;
    precond(a.minv, a.precond, p, rr.x, rr.y, rr.z, z0, z1, z2);
    a.z[p] = make_double4(z0, z1, z2, 0.0);
    t_rz = term3(rr.x, rr.y, rr.z, z0, z1, z2);
    t_rr = term3(rr.x, rr.y, rr.z, rr.x, rr.y, rr.z);
  }
  const double v0 = tile_sum(t_rz, lds4);
  const double v1 = tile_sum(t_rr, lds4);
  if (threadIdx.x == 0) {
    a.part[2 * (int64_t)ntiles + tile] = v0;
    a.part[3 * (int64_t)ntiles + tile] = v1;
  }
}

// alpha of the specimens with more than PD_ALPHA_FOLD_TILES tiles, one
// workgroup each (the same fold and quotient as k_cg_update's)
__global__ __launch_bounds__(PD_BLOCK) void k_cg_alpha(CgArgs a, int ntiles) {
  pd_cg_state* st = a.st + blockIdx.x;
  if (st->done) return;
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  if (nt <= PD_ALPHA_FOLD_TILES) return;
  const double yp = fold_partials(a.part + t0, nt, lds4);
  if (threadIdx.x == 0) {
    st->yp = yp;
    st->alpha = st->rz / yp;
  }
}

// beta = rz' / rz, iteration count and the monitor test ahead of the next
// iteration (||r|| <= tol converged, iter >= max_iter failed), for every
// specimen whose update ran in iteration `it`
__global__ __launch_bounds__(PD_BLOCK) void k_cg_beta(CgArgs a, int ntiles, int it) {
  if (a.st[blockIdx.x].upd_it != it) return;
  __shared__ double lds4[4];
  int t0, nt;
  spec_range(a.batch, ntiles, t0, nt);
  const double rz = fold_partials(a.part + 2 * (int64_t)ntiles + t0, nt, lds4);
  const double rr = fold_partials(a.part + 3 * (int64_t)ntiles + t0, nt, lds4);
  if (threadIdx.x == 0) {
    pd_cg_state* st = a.st + blockIdx.x;
    const double rz_old = st->rz;
    st->rz = rz;
    st->beta = rz / rz_old;
    st->iter = st->iter + 1;
    st->rr = rr;
    const double rnorm = sqrt(rr);
    st->rnorm = rnorm;
    if (rnorm <= st->tol) {
      st->done = 1;
      st->converged = 1;
    } else if (st->iter >= st->max_iter) {
      st->done = 1;
    }
  }
}

// The end of iteration `it` for every specimen whose update ran in it:
// x += alpha p, and, unless k_cg_beta found the solve ended, p = z + beta p.
__global__ __launch_bounds__(PD_BLOCK) void k_cg_dir(CgArgs a, int it) {
  const int tile = xcd_tile(blockIdx.x, gridDim.x);
  const pd_cg_state* st = a.st + spec_of(a.batch, tile);
  if (st->upd_it != it) return;  // the solve had ended before this iteration
  const int p = tile * PD_BLOCK + threadIdx.x;
  if (p >= a.A.n) return;
  // matrix-free (A.rec set): p lives in the gather record alone
  double4 pp = a.A.rec ? rec_vec(a.A.rec, p) : a.p[p];
  double4 xx = a.x[p];
  const double alpha = st->alpha;
  xx.x = alpha * pp.x + xx.x;
  xx.y = alpha * pp.y + xx.y;
  xx.z = alpha * pp.z + xx.z;
  a.x[p] = xx;
  if (st->done) return;
  const double beta = st->beta;
  const double4 zz = a.z[p];
  pp.x = zz.x + beta * pp.x;
  pp.y = zz.y + beta * pp.y;
  pp.z = zz.z + beta * pp.z;
  if (a.A.rec)
    rec_p(a.A.rec, p, pp.x, pp.y, pp.z);
  else
    a.p[p] = pp;
}

static inline int tiles_of(int n) { return (n + PD_BLOCK - 1) / PD_BLOCK; }

void launch_cg_init(const CgArgs& a, hipStream_t s) {
  if (!a.A.n) return;
  const int g = tiles_of(a.A.n);
  if (a.A.mf) {
    hipLaunchKernelGGL(k_cg_init<0>, dim3(g), dim3(PD_BLOCK), 0, s, a);
  } else {
    switch (a.A.unroll) {
      case 2: hipLaunchKernelGGL(k_cg_init<2>, dim3(g), dim3(PD_BLOCK), 0, s, a); break;
      case 8: hipLaunchKernelGGL(k_cg_init<8>, dim3(g), dim3(PD_BLOCK), 0, s, a); break;
      default: hipLaunchKernelGGL(k_cg_init<4>, dim3(g), dim3(PD_BLOCK), 0, s, a);
    }
  }
  hipLaunchKernelGGL(k_cg_init_fold<false>, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g,
                     0.0, 0);
}
void launch_cg_init_store(const CgArgs& a, hipStream_t s) {
  if (!a.A.n) return;
  const int g = tiles_of(a.A.n);
  hipLaunchKernelGGL(HIP_KERNEL_NAME(k_cg_init<0, true>), dim3(g), dim3(PD_BLOCK), 0, s, a);
  hipLaunchKernelGGL(k_cg_init_fold<false>, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g,
                     0.0, 0);
}
void launch_cg_init_fold(const CgArgs& a, double rtol, int32_t max_iter, hipStream_t s) {
  if (a.A.n)
    hipLaunchKernelGGL(k_cg_init_fold<true>, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a,
                       tiles_of(a.A.n), rtol, max_iter);
  else
    launch_cg_setup(a.st, a.batch.nspec, rtol, max_iter, s);
}
void launch_cg_spmv(const CgArgs& a, hipStream_t s) {
  if (!a.A.n) return;
  const int g = tiles_of(a.A.n);
  if (a.A.mf) {
    if (a.A.rec && a.A.range_bad)  // after an assembly: the record and the range flag
      hipLaunchKernelGGL(k_cg_spmv_mf, dim3(g), dim3(PD_BLOCK), 0, s, a);
    else
      hipLaunchKernelGGL(k_cg_spmv<0>, dim3(g), dim3(PD_BLOCK), 0, s, a);
    return;
  }
  switch (a.A.unroll) {
    case 2: hipLaunchKernelGGL(k_cg_spmv<2>, dim3(g), dim3(PD_BLOCK), 0, s, a); break;
    case 8: hipLaunchKernelGGL(k_cg_spmv<8>, dim3(g), dim3(PD_BLOCK), 0, s, a); break;
    default: hipLaunchKernelGGL(k_cg_spmv<4>, dim3(g), dim3(PD_BLOCK), 0, s, a);
  }
}
void launch_cg_spmv_store(const CgArgs& a, hipStream_t s) {
  if (!a.A.n) return;
  hipLaunchKernelGGL(k_cg_spmv_store, dim3(tiles_of(a.A.n)), dim3(PD_BLOCK), 0, s, a);
}
void launch_write_blocks(const SpmvView& A, double* vals, hipStream_t s) {
  if (!A.n) return;
  hipLaunchKernelGGL(k_write_blocks, dim3(tiles_of(A.n)), dim3(PD_BLOCK), 0, s, A, vals);
}
void launch_cg_update(const CgArgs& a, int it, hipStream_t s) {
  if (!a.A.n) return;
  const int g = tiles_of(a.A.n);
  if (g > PD_ALPHA_FOLD_TILES)  // some specimen may exceed the threshold
    hipLaunchKernelGGL(k_cg_alpha, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g);
  hipLaunchKernelGGL(k_cg_update, dim3(g), dim3(PD_BLOCK), 0, s, a, it);
}
void launch_cg_dir(const CgArgs& a, int it, hipStream_t s) {
  if (!a.A.n) return;
  const int g = tiles_of(a.A.n);
  hipLaunchKernelGGL(k_cg_beta, dim3(a.batch.nspec), dim3(PD_BLOCK), 0, s, a, g, it);
  hipLaunchKernelGGL(k_cg_dir, dim3(g), dim3(PD_BLOCK), 0, s, a, it);
}

// ---------------------------------------------------------------------------
// Position update: core/implicit_euler.cuh:90-176, core/newmark_beta.cuh:
// 89-199 (velocity then position with the per-axis wall clamp).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void clamp_axis(double& pos, double& vel, double lo, double hi,
                                           double rest) {
  if (pos > hi) {
    vel *= -rest;
    pos = hi - rest * (pos - hi);
  }
  if (pos < lo) {
    vel *= -rest;
    pos = lo + rest * (lo - pos);
  }
}

struct UpdArgs {
  int n;
  double4* x;
  double4* v;
  const double4* sol;
  const int32_t* act;
  int integrator;
  double dt;
  int walls;
  double lo[3];
  double hi[3];
  double rest;
  const pd_step_consts* consts;
};

__global__ __launch_bounds__(PD_BLOCK) void k_update(UpdArgs a) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (p >= a.n || a.act[p] == PD_INVISIBLE) return;
  if (a.consts) a.dt = a.consts->dt;
  const double4 s = a.sol[p];
  const double4 xo = a.x[p];
  double nv0, nv1, nv2, np0, np1, np2;
  if (a.integrator == 1) {
    const double4 vo = a.v[p];
    const double c = 2.0 / a.dt;
    nv0 = c * s.x - vo.x;
    nv1 = c * s.y - vo.y;
    nv2 = c * s.z - vo.z;
    np0 = xo.x + s.x;
    np1 = xo.y + s.y;
    np2 = xo.z + s.z;
  } else {
    nv0 = s.x;
    nv1 = s.y;
    nv2 = s.z;
    np0 = xo.x + nv0 * a.dt;
    np1 = xo.y + nv1 * a.dt;
    np2 = xo.z + nv2 * a.dt;
  }
  if (a.walls) {
    clamp_axis(np0, nv0, a.lo[0], a.hi[0], a.rest);
    clamp_axis(np1, nv1, a.lo[1], a.hi[1], a.rest);
    clamp_axis(np2, nv2, a.lo[2], a.hi[2], a.rest);
  }
  a.x[p] = make_double4(np0, np1, np2, 0.0);
  a.v[p] = make_double4(nv0, nv1, nv2, 0.0);
}

void launch_update(int32_t n, double4* x, double4* v, const double4* sol,
                   const int32_t* act, int32_t integrator, double dt, int32_t walls,
                   const double* lo, const double* hi, double restitution,
                   const pd_step_consts* consts, hipStream_t s) {
  UpdArgs a;
  a.consts = consts;
  a.n = n;
  a.x = x;
  a.v = v;
  a.sol = sol;
  a.act = act;
  a.integrator = integrator;
  a.dt = dt;
  a.walls = walls;
  for (int k = 0; k < 3; ++k) {
    a.lo[k] = lo[k];
    a.hi[k] = hi[k];
  }
  a.rest = restitution;
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_update, dim3(nb), dim3(PD_BLOCK), 0, s, a);
}

// d0 of every slot below the row's capacity; *bad (cleared first by
// k_clear_word) set if one lies outside range_ok_len
__global__ __launch_bounds__(PD_BLOCK) void k_bond_d0(int n, const int64_t* slice_off,
                                                      const int32_t* cap, const int32_t* col,
                                                      const double4* x0, double* d0,
                                                      unsigned int* bad) {
  const int p = blockIdx.x * PD_BLOCK + threadIdx.x;
  bool out = false;
  if (p < n) {
    const int64_t base = slice_off[p >> 6];
    const int lane = p & 63;
    const double4 op = x0[p];
    for (int k = 0; k < cap[p]; ++k) {
      const int64_t i = (base + k) * 64 + lane;
      const double l = ref_len(op, x0[col[i]]);
      d0[i] = l;
      out = out || !range_ok_len(l);
    }
  }
  if (__any(out) && (threadIdx.x & 63) == 0) atomicOr(bad, 1u);
}

__global__ void k_clear_word(unsigned int* w) {
  if (threadIdx.x == 0) *w = 0u;
}

void launch_bond_d0(int32_t n, const int64_t* slice_off, const int32_t* cap, const int32_t* col,
                    const double4* x0, double* d0, unsigned int* bad, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, s, bad);
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb)
    hipLaunchKernelGGL(k_bond_d0, dim3(nb), dim3(PD_BLOCK), 0, s, n, slice_off, cap, col, x0, d0,
                       bad);
}

__global__ __launch_bounds__(PD_BLOCK) void k_copy(double* dst, const double* src, int n) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

void launch_copy(double* dst, const double* src, int32_t n, hipStream_t s) {
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(PD_BLOCK), 0, s, dst, src, n);
}

// ---------------------------------------------------------------------------
// Contact: core/simulator.cuh:679-766 collidePDParticles on the uniform grid
// (27 cells in z, y, x order; members in ascending index, as the stable
// sort leaves them), between particles of different bodies or, with
// all_pairs, between any two particles.
// ---------------------------------------------------------------------------
struct ContactArgs {
  int n;
  const double4* x;
  const int32_t* body;
  const int32_t* act;
  GridView g;
  const int32_t* cand;
  int m;
  double Rc, kc, vol;
  int all_pairs;
  double4* fc;
};

__global__ __launch_bounds__(PD_BLOCK) void k_zero4(double4* v, int n) {
  const int i = blockIdx.x * PD_BLOCK + threadIdx.x;
  if (i < n) v[i] = make_double4(0.0, 0.0, 0.0, 0.0);
}

// One wave per candidate p.  The members of its 27 cells are numbered in
// the serial order (cells in z, y, x order, members ascending); each lane
// evaluates one member's term at a time, and the terms are added into f in
// that order (broadcast lane by lane), so f equals the serial sum bit for
// bit.  The member loads and square roots of a candidate run in parallel
// instead of one after another.
__global__ __launch_bounds__(PD_BLOCK) void k_contact(ContactArgs a) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * (PD_BLOCK / 64) + (threadIdx.x >> 6);
  if (i >= a.m) return;  // whole waves
  const int p = a.cand ? a.cand[i] : i;
  if (a.act[p] == PD_INVISIBLE) return;  // fc[p] stays zero
  const double4 pp = a.x[p];
  const int bp = a.body[p];
  const int cx = (int)floor((pp.x - a.g.mn[0]) / a.g.cell) - a.g.c0[0];
  const int cy = (int)floor((pp.y - a.g.mn[1]) / a.g.cell) - a.g.c0[1];
  const int cz = (int)floor((pp.z - a.g.mn[2]) / a.g.cell) - a.g.c0[2];
  // lane j < 27: cell j's member range (dx fastest, then dy, then dz)
  int st = 0, cnt = 0;
  if (lane < 27) {
    const int ix = cx + lane % 3 - 1, iy = cy + (lane / 3) % 3 - 1, iz = cz + lane / 9 - 1;
    if (ix >= 0 && iy >= 0 && iz >= 0 && ix < a.g.dims[0] && iy < a.g.dims[1] &&
        iz < a.g.dims[2]) {
      const int64_t cell = ((int64_t)iz * a.g.dims[1] + iy) * a.g.dims[0] + ix;
      st = a.g.cell_start[cell];
      cnt = a.g.cell_end[cell] - st;
    }
  }
  int incl = cnt;  // inclusive prefix sum of the counts over the lanes
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int v = __shfl_up(incl, d, 64);
    if (lane >= d) incl += v;
  }
  const int excl = incl - cnt;
  const int total = __shfl(incl, 63, 64);
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  for (int base = 0; base < total; base += 64) {
    const int t = base + lane;
    int m = -1;
    for (int j = 0; j < 27; ++j) {
      const int ej = __shfl(excl, j, 64), cj = __shfl(cnt, j, 64), sj = __shfl(st, j, 64);
      if (t >= ej && t < ej + cj) m = sj + (t - ej);
    }
    int use = 0;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    if (t < total && m >= 0) {
      const int q = a.g.sorted_idx[m];
      if (q != p && (a.all_pairs || a.body[q] != bp) && a.act[q] != PD_INVISIBLE) {
        const double4 xq = a.x[q];
        const double q0 = pp.x - xq.x, q1 = pp.y - xq.y, q2 = pp.z - xq.z;
        const double r = sqrt(q0 * q0 + q1 * q1 + q2 * q2);
        if (r < a.Rc && r != 0.0) {
          const double howfar = (r - a.Rc);
          const double mag = a.kc * (howfar * howfar);
          const double coef = mag / a.vol;
          t0 = coef * (q0 / r);
          t1 = coef * (q1 / r);
          t2 = coef * (q2 / r);
          use = 1;
        }
      }
    }
    const int nk = min(64, total - base);
    for (int k = 0; k < nk; ++k) {  // the serial order
      const int u = __shfl(use, k, 64);
      const double a0 = __shfl(t0, k, 64), a1 = __shfl(t1, k, 64), a2 = __shfl(t2, k, 64);
      if (u) {
        f0 += a0;
        f1 += a1;
        f2 += a2;
      }
    }
  }
  if (lane == 0) a.fc[p] = make_double4(f0, f1, f2, 0.0);
}

void launch_contact(int32_t n, const double4* x, const int32_t* body, const int32_t* act,
                    const GridView& g, const int32_t* cand, int32_t m, double Rc, double kc,
                    double vol, int all_pairs, double4* fc, hipStream_t s, bool fc_zeroed) {
  const int nb = (n + PD_BLOCK - 1) / PD_BLOCK;
  if (nb && !fc_zeroed) hipLaunchKernelGGL(k_zero4, dim3(nb), dim3(PD_BLOCK), 0, s, fc, n);
  if (m <= 0) return;
  ContactArgs a;
  a.n = n;
  a.x = x;
  a.body = body;
  a.act = act;
  a.g = g;
  a.cand = cand;
  a.m = m;
  a.Rc = Rc;
  a.kc = kc;
  a.vol = vol;
  a.all_pairs = all_pairs;
  a.fc = fc;
  const int per = PD_BLOCK / 64;  // candidates per workgroup, one wave each
  const int mb = (m + per - 1) / per;
  hipLaunchKernelGGL(k_contact, dim3(mb), dim3(PD_BLOCK), 0, s, a);
}

}  // namespace pd
