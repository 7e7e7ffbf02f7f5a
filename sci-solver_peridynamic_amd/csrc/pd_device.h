// pd_device.h -- device-side layout and helpers shared by the HIP kernels.
//
// HBM layout (one context, N particles, S = ceil(N/64) slices):
//   * particle vectors: double4 per particle (32 B, w lane 0) so that the
//     gathered x_q of a bond is one aligned 32-byte segment;
//   * bonds: sliced ELL with 64 rows per slice (one wave64 lane per row).
//     Slice s owns slot rows [slice_off[s], slice_off[s+1]); the slot k of
//     row p = 64 s + lane lives at (slice_off[s] + k) * 64 + lane, so the
//     64 lanes of a wave touch 64 consecutive words per slot;
//   * matrix values: the 3x3 blocks are symmetric bit for bit, so six
//     values (a00 a01 a02 a11 a12 a22) are stored.  FP64: in three planes
//     of pairs, so a lane reads 16 bytes per load: value c of slot k of row
//     p at ((slice_off[s] + k) * 3 + c / 2) * 128 + 2 * lane + c % 2.
//     Float copy (val32): 384 floats per slot row, a plane of quads
//     (a00 a01 a02 a11) then a plane of pairs (a12 a22), so a lane reads
//     one 16-byte and one 8-byte word: at (slice_off[s] + k) * 384, the
//     quad at 4 * lane and the pair at 256 + 2 * lane;
//   * diagonal and block-Jacobi inverse blocks: (s * 6 + c) * 64 + lane;
//   * mixed precision keeps float copies in the same positions (val32,
//     diag32, minv32) and float4 vectors (16 B) for the inner solve.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define PD_WAVE 64
#define PD_BLOCK 256  // 4 waves = 4 slices = one reduction tile
#define PD_INVISIBLE 3

struct pd_mat_dev {
  double horizon, radius, c_bond, mass;
};

// Scalars of one CG solve, resident in device memory.
struct pd_cg_state {
  double rz;     // <r, z>
  double yp;     // <A p, p>
  double alpha;
  double beta;
  double rr;     // <r, r>
  double bnorm;  // ||S b||
  double tol;    // rtol * bnorm
  double rnorm;  // sqrt(rr)
  double rtol;
  int32_t iter;
  int32_t done;
  int32_t converged;
  int32_t max_iter;
  int32_t upd_it;  // the iteration whose update ran last (-1: none yet)
  uint32_t pad[3];
};

// Scalars of one mixed-precision solve (iterative refinement around a float
// PCG).  phase: 0 residual due, 1 inner iterations, 2 inner finished
// (correction pending, then residual), 3 done.  Grid kernels only read
// phase; the one-workgroup fold kernels change it.
struct pd_mx_state {
  double rz;         // <r32, z32> (FP64 sum of widened floats)
  double rr32;       // <r32, r32>
  double alpha;
  double beta;
  double rnorm;      // ||S(b - A x)||, FP64
  double bnorm;      // ||S b||
  double tol;        // rtol * bnorm
  double inner_tol;  // max(inner_rtol * rnorm, tol / 2)
  double rtol;
  double inner_rtol;
  int32_t iter;        // inner iterations over the whole solve
  int32_t inner_iter;  // of the current refinement
  int32_t refinements;
  int32_t phase;
  int32_t converged;
  int32_t first;
  int32_t max_iter;
  int32_t inner_max;
  int32_t max_refine;
  int32_t residuals;  // FP64 residual evaluations so far
  int32_t pad[2];
};

// Step scalars read from device memory by captured (graph) steps, so one
// graph serves every step.
struct pd_step_consts {
  double dt;
  double dt2;
  double clock;
  double gravity;
};

// Per-step counters written by the assembly kernel.
struct pd_step_counters {
  unsigned long long broken;
  unsigned long long live;
  // bit 0: some position, reference length, bond stiffness, horizon or
  // radius lies outside the ranges of the shared-reciprocal quotients and
  // of sqrt_nr (set ahead of the assembly); bit 1: a live bond has zero
  // length (set by the assembly).  Nonzero: the matrix-free operator
  // divides one quotient at a time and takes IEEE square roots.
  unsigned int range_bad;
  unsigned int pad;
};

// The CG updates form alpha = rz / <y, p> in every workgroup from the
// specimen's <y, p> tile partials (one read of the partial row per
// workgroup, hidden behind the row loads); a specimen of more tiles than
// this has alpha folded once, by one workgroup, ahead of the update, since
// the per-workgroup reads grow with the square of its tile count.
#define PD_ALPHA_FOLD_TILES 8192

__device__ __forceinline__ double4 ld4(const double4* p) { return *p; }

// Blocks of the workgroup grid are handed to the 8 XCDs round-robin; this
// maps the hardware order onto contiguous tile ranges per XCD so that the
// tiles sharing an L2 are neighbours in the particle order.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int xcd = b & 7;
  const int idx = b >> 3;
  const int q = nb >> 3;
  const int r = nb & 7;
  return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Sum of a double over the 64 lanes by the xor butterfly m = 32..1; every
// lane ends with the same value (the order fixed by the oracle's orc_dot).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 64);
  return v;
}

// Tile reduction of one value per thread of a 256-thread workgroup:
// (w0 + w1) + (w2 + w3) of the four wave sums.  All threads must call it.
__device__ __forceinline__ double tile_sum(double v, double* lds4) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) lds4[w] = v;
  __syncthreads();
  const double r = (lds4[0] + lds4[1]) + (lds4[2] + lds4[3]);
  __syncthreads();
  return r;
}

// Final fold of one quantity over the tiles by one 256-thread workgroup:
// thread t sums part[t], part[t + 256], ... left to right from 0.0, then
// tile_sum (the oracle's final_combine).
__device__ __forceinline__ double fold_partials(const double* part, int ntiles, double* lds4) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < ntiles; k += PD_BLOCK) acc = acc + part[k];
  return tile_sum(acc, lds4);
}
