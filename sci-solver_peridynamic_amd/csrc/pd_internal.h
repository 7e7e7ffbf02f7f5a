// pd_internal.h -- launchers shared between the kernel files and the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pd_device.h"

namespace pd {

// Bond topology in the sliced layout (see pd_device.h).
struct BondView {
  int32_t n;
  int32_t nslices;
  const int64_t* slice_off;  // [nslices + 1], in slot rows
  int32_t* top;              // [n] last live slot (-1: none)
  int32_t* col;              // [slots * 64]
};

struct AsmArgs {
  int32_t n;
  const int64_t* slice_off;
  int32_t* top;
  int32_t* col;
  double* val;
  double* diag;
  double* minv;
  float* diag32;  // mixed precision: float diagonal and inverse beside the FP64 ones
  float* minv32;
  int32_t mixed;
  const double4* x;
  const double4* v;
  const double4* fc;
  const double4* vbc;
  double4* f;
  double4* b;
  double4* sol;
  const double* s;
  double* s_new;
  const double* s00;
  const int32_t* act;
  const int32_t* body;
  const int32_t* bc;
  const pd_mat_dev* mat;  // device copy of the body table
  double dt;
  double dt2;
  double clock;
  double gravity;
  int32_t integrator;
  int32_t contact;
  int32_t track_gap;
  const pd_step_consts* consts;  // non-null: dt, dt2, clock, gravity from here
  float2* bond_stats;            // non-null: (mean, max) stretch of the live bonds
  double* d0;            // per-slot reference lengths: read, and swapped with col
  // fuse = 1: the solve's first residual r = S(b - A xs), each row's bond
  // terms summed first and the diagonal term last, with the solver's start
  // and its tile partials in part (Dirichlet rows of xs must hold their
  // prescribed values already: k_asm_prep).  z = r (pc = 0) or
  // the block-Jacobi inverse just formed times r (pc = 1).
  //   FP64:  r64 = r, p64 = z (p64 null: a matrix-free solve, whose p lives
  //          in the record rec alone); partials <r,z>, <r,r>, <Sb,Sb>
  //   mixed: r32 = float(r), p32 = z, d32 = 0 (A32 with float bond blocks);
  //          partials <r,r>, <r32,z>, <r32,r32>, <Sb,Sb>
  int32_t fuse;
  int32_t pc;
  const double4* xs;
  double4* r64;
  double4* p64;
  float4* r32;
  float4* p32;
  float4* d32;
  double* part;
  double* gap_val;   // [n] when track_gap
  int32_t* gap_q;    // [n] when track_gap
  pd_step_counters* counters;  // range_bad (set by k_asm_prep) selects SHARED
  double* rec;  // non-null: write the p part of the matrix-free gather record
  // per particle (x, y, z, s, xs.x, xs.y, xs.z, 0), written by k_asm_prep
  const double* arec;
};

// The assembly record and the quotients' range flag ahead of the assembly
// (k_asm_prep); with fuse, the Dirichlet rows' start values first.
struct AsmPrepArgs {
  int32_t n;
  int32_t fuse;
  int32_t integrator;
  int32_t c_bad;  // a body's bond stiffness, horizon or radius outside range_ok_stiff
  double dt;
  const pd_step_consts* consts;  // non-null: dt from here
  const double4* x;
  const double* s;
  double4* sol;
  const int32_t* bc;
  const int32_t* act;
  const double4* vbc;
  const unsigned int* d0_bad;  // k_bond_d0's range flag
  double* arec;                // [n][8]
  double* rec;                 // non-null: the matrix-free record's position part
  pd_step_counters* counters;
};

struct SpmvView {
  int32_t n;
  const int64_t* slice_off;
  const int32_t* top;
  const int32_t* col;
  const double* val;
  const double* diag;
  int32_t unroll;  // slots per trip of the SpMV row loop: 2, 4 or 8
  // mf = 1: matrix-free operator.  val is not read; each live bond's block
  // is recomputed from the step's positions (pos), the slot's reference
  // length (d0), the row's material and dt2 (from consts when non-null)
  // exactly as the assembly forms it.
  int32_t mf;
  const double4* pos;
  const int32_t* body;
  const pd_mat_dev* mat;
  double dt2;
  const pd_step_consts* consts;
  const double* d0;  // mf: reference bond length per slot
  // mf, both set only between an assembly and the next position change:
  // range_bad non-null: the assembly's range flag (0: the bond quotients
  // may share reciprocals); rec non-null: per particle the position and
  // the CG direction p side by side (6 doubles), kept by the assembly and
  // every kernel that writes p, so one gather serves both
  const unsigned int* range_bad;
  double* rec;
};

// Independent specimens solved together: specimen k owns the 256-row
// tiles [spec_tile0[k], spec_tile0[k] + spec_ntiles[k]); tile_spec maps a
// tile to its specimen.  Every solver scalar is per specimen, every sum is
// folded per specimen.  Null tables: one specimen over all tiles.
struct Batch {
  const int32_t* tile_spec;
  const int32_t* spec_tile0;
  const int32_t* spec_ntiles;
  int32_t nspec;
};

__device__ __forceinline__ int spec_of(const Batch& b, int tile) {
  return b.tile_spec ? b.tile_spec[tile] : 0;
}
// tile range of specimen s
__device__ __forceinline__ void spec_range_of(const Batch& b, int s, int ntiles, int& t0,
                                              int& nt) {
  if (b.spec_tile0) {
    t0 = b.spec_tile0[s];
    nt = b.spec_ntiles[s];
  } else {
    t0 = 0;
    nt = ntiles;
  }
}
// tile range of the specimen folded by this workgroup (blockIdx.x)
__device__ __forceinline__ void spec_range(const Batch& b, int ntiles, int& t0, int& nt) {
  spec_range_of(b, blockIdx.x, ntiles, t0, nt);
}

struct CgArgs {
  SpmvView A;
  Batch batch;
  const double* minv;
  const int32_t* bc;
  const double4* b;
  double4* x;
  double4* r;
  double4* z;
  double4* p;
  double4* y;
  double* part;  // [4 * ntiles]
  pd_cg_state* st;  // [batch.nspec]
  int32_t precond;
  // stored-block FP64 steps whose assembly left the blocks to the first
  // product: that product (launch_cg_spmv_store) recomputes every live
  // bond's block from the gather record rec (positions and p) and writes it
  // to vals, which the later products stream
  double* vals;
  const double* rec;
};

// Mixed-precision solve: float matrix copy, float inner vectors, FP64
// solution and right-hand side.
struct MxArgs {
  int32_t n;
  Batch batch;
  const int64_t* slice_off;
  const int32_t* top;
  const int32_t* col;
  const float* val32;
  const double* diag;
  const float* diag32;
  const float* minv32;
  const int32_t* bc;
  const double4* b;
  double4* x;
  float4* r;
  float4* z;
  float4* p;
  float4* y;
  float4* d;
  double* part;  // [4 * ntiles]
  pd_mx_state* st;  // [batch.nspec]
  int32_t precond;
  int32_t unroll;  // slots per trip of the SpMV row loops: 2, 4 or 8
  // the FP64 system of the residual: its bond blocks recomputed from the
  // positions as the assembly forms them (matrix-free), FP64 diagonal.
  // A64.rec, set between a mixed assembly and the next position change,
  // holds per particle the position and x side by side (the apply and
  // launch_mx_rec_x write x)
  SpmvView A64;
};

// Learned block preconditioner: a 16-64-64-8 MLP (BF16 weights, FP32
// accumulation on MFMA) maps per-particle features to S_p = L L^T + eps I,
// and M_p = C_p^{-T} S_p C_p^{-1} with D_p = C_p C_p^T (S = I is block Jacobi).
#define PC_NIN 16
#define PC_NH 64
#define PC_NOUT 8
// flat parameter vector: W1[NH][NIN] b1[NH] W2[NH][NH] b2[NH] W3[NOUT][NH]
// b3[NOUT] fmean[NIN] fistd[NIN] eps
#define PC_OFF_W1 0
#define PC_OFF_B1 (PC_OFF_W1 + PC_NH * PC_NIN)
#define PC_OFF_W2 (PC_OFF_B1 + PC_NH)
#define PC_OFF_B2 (PC_OFF_W2 + PC_NH * PC_NH)
#define PC_OFF_W3 (PC_OFF_B2 + PC_NH)
#define PC_OFF_B3 (PC_OFF_W3 + PC_NOUT * PC_NH)
#define PC_OFF_MEAN (PC_OFF_B3 + PC_NOUT)
#define PC_OFF_ISTD (PC_OFF_MEAN + PC_NIN)
#define PC_OFF_EPS (PC_OFF_ISTD + PC_NIN)
#define PC_NPARAM (PC_OFF_EPS + 1)
#define PC_NWEIGHT PC_OFF_MEAN  // trainable entries (weights and biases)

struct PcNet {
  const __bf16* w;     // [PC_NWEIGHT] bf16 copy (biases unused in this copy)
  const float* p;      // [PC_NPARAM] FP32 parameters
};

// per-particle raw features (float[n][PC_NIN]) of the assembled system
struct PcFeatArgs {
  int32_t n;
  const int64_t* slice_off;
  const int32_t* top;
  const int32_t* cap;
  const int32_t* col;
  const double4* x;
  const double* diag;  // FP64 planes
  const float2* bond_stats;
  const int32_t* act;
  const int32_t* body;
  const pd_mat_dev* mat;
  float* feat;
};
void launch_pc_features(const PcFeatArgs& a, hipStream_t s);
// M_p from the network for every visible row: minv (FP64 planes) and, when
// minv32 is non-null, its float copy
void launch_pc_infer(int32_t n, const float* feat, const PcNet& net, const double* diag,
                     const int32_t* act, double* minv, float* minv32, hipStream_t s);
// launch_pc_infer and the solve's start with the learned blocks in one
// pass, after an assembly that formed the first residual (fuse): z = M r,
// p = z, the <r, z> tile partials
struct PcInferStartArgs {
  int32_t n;
  int32_t mixed;
  const float* feat;
  PcNet net;
  const double* diag;
  const int32_t* act;
  double* minv;          // the learned blocks (FP64)
  float* minv32;         // and rounded to float (mixed 1)
  const double4* r64;
  double4* p64;          // null: a matrix-free solve (p in rec alone)
  double* rec;           // FP64: the gather record's p part, when in use
  const float4* r32;
  float4* p32;
  double* part;          // the solve's tile partials
};
void launch_pc_infer_start(const PcInferStartArgs& a, hipStream_t s);

// training samples: feat[NIN], D (6, float), v = (A e)_p (3), e_p (3)
#define PC_SAMPLE 28
struct PcSampleArgs {
  int32_t n;
  const float* feat;
  const double* diag;
  const double4* Ae;  // A e
  const double4* e;
  const int32_t* act;
  const int32_t* bc;
  uint64_t seed;
  uint32_t keep;  // stride: rows p with p % keep == (hash of seed) % keep
  float* out;     // [capacity][PC_SAMPLE], sample p / keep
  int64_t capacity;
};
void launch_pc_probe(int32_t n, uint64_t seed, double4* e, hipStream_t s);
void launch_pc_samples(const PcSampleArgs& a, hipStream_t s);
// one optimisation step on `batch` samples drawn with replacement
struct PcTrainArgs {
  const float* samples;
  int64_t nsamples;
  int32_t batch;  // multiple of 32
  uint64_t seed;
  int32_t step;
  const __bf16* w;  // bf16 weights
  const float* p;   // FP32 parameters (biases, normalisation, eps)
  float* partial;   // [batch / 32][PC_NWEIGHT + 1] per-wave gradients + loss
};
void launch_pc_train(const PcTrainArgs& a, hipStream_t s);
// fold the partial gradients, Adam update of p (weights and biases), refresh w
void launch_pc_adam(const float* partial, int32_t nwaves, float* p, float* m, float* v,
                    __bf16* w, int32_t step, float lr, float* loss_out, hipStream_t s);
void launch_pc_to_bf16(const float* p, __bf16* w, hipStream_t s);
// per-feature mean and 1/std over the samples' raw features into p
void launch_pc_norm(const float* samples, int64_t nsamples, float* p, hipStream_t s);

#define PD_MAX_ZONE_BODIES 8

struct GridView {
  double mn[3];
  double cell;
  int32_t c0[3];  // cell of the sub-grid's first cell in the full grid
  int32_t dims[3];
  const uint32_t* sorted_key;
  const int32_t* sorted_idx;
  const int32_t* cell_start;
  const int32_t* cell_end;
};

void launch_assemble(const AsmArgs& a, hipStream_t s);
void launch_asm_prep(const AsmPrepArgs& a, hipStream_t s);
void launch_gap_reduce(const double* gap_val, const int32_t* gap_q, int32_t n,
                       double* part_val, int32_t* part_p, int32_t* part_q,
                       double* out_val, int32_t* out_pq, hipStream_t s);
void launch_cg_setup(pd_cg_state* st, int32_t nspec, double rtol, int32_t max_iter,
                     hipStream_t s);
void launch_cg_init(const CgArgs& a, hipStream_t s);
// the solver state set up (launch_cg_setup's) and the fold of the start's
// partials, after an assembly that produced them
void launch_cg_init_fold(const CgArgs& a, double rtol, int32_t max_iter, hipStream_t s);
void launch_cg_spmv(const CgArgs& a, hipStream_t s);
// the first CG product of a step whose assembly did not write the blocks:
// y = A p with the blocks recomputed and written to a.vals (for every row,
// converged specimens included), <y,p> partials as launch_cg_spmv
void launch_cg_spmv_store(const CgArgs& a, hipStream_t s);
// launch_cg_init when the assembly left the blocks: they are recomputed
// from the positions, written to a.vals and multiplied by x0
void launch_cg_init_store(const CgArgs& a, hipStream_t s);
// the blocks of the current bond lists written to vals from the positions
// (a step's deferred blocks when no product wrote them)
void launch_write_blocks(const SpmvView& A, double* vals, hipStream_t s);
// iteration `it` (0-based, counted from the solve's start) of the CG loop
void launch_cg_update(const CgArgs& a, int it, hipStream_t s);
void launch_cg_dir(const CgArgs& a, int it, hipStream_t s);
void launch_spmv_plain(const SpmvView& A, const double4* x, double4* y, hipStream_t s);
// pd_mixed.hip
void launch_mx_setup(pd_mx_state* st, int32_t nspec, double rtol, int32_t max_iter,
                     double inner_rtol, int32_t inner_max, int32_t max_refine, hipStream_t s);
void launch_mx_resid(const MxArgs& a, hipStream_t s);  // residual + its fold
// x into the gather record ahead of a residual that no apply preceded
// (a no-op unless A64.rec is set)
void launch_mx_rec_x(const MxArgs& a, hipStream_t s);
// the fold alone, after an assembly that produced the residual partials
void launch_mx_resid_fold(const MxArgs& a, hipStream_t s);
void launch_mx_spmv(const MxArgs& a, hipStream_t s);
// the first inner product after a mixed assembly, which leaves the float
// blocks: they are recomputed from the positions, written to vals (every
// row) and multiplied
void launch_mx_spmv_store(const MxArgs& a, float* vals, hipStream_t s);
// the float blocks of the current bond lists written to vals from the
// positions (a step's deferred float blocks when no product wrote them)
void launch_write_blocks32(const SpmvView& A, float* vals, hipStream_t s);
void launch_mx_update(const MxArgs& a, hipStream_t s);  // alpha, update, beta
void launch_mx_dir(const MxArgs& a, hipStream_t s);
void launch_mx_apply(const MxArgs& a, hipStream_t s);
// y = A32 x with FP64 arithmetic (float bond blocks widened; pd_spmv after a
// mixed assembly), and y32 = A32 x32
void launch_spmv_wide(const MxArgs& a, const double4* x, double4* y, hipStream_t s);
void launch_spmv_f32(const MxArgs& a, const float4* x, float4* y, hipStream_t s);
// consts non-null: dt from there
void launch_update(int32_t n, double4* x, double4* v, const double4* sol,
                   const int32_t* act, int32_t integrator, double dt, int32_t walls,
                   const double* lo, const double* hi, double restitution,
                   const pd_step_consts* consts, hipStream_t s);
void launch_copy(double* dst, const double* src, int32_t n, hipStream_t s);
// Compares shared-reciprocal quotients with IEEE division on `samples`
// random operand pairs; *bad counts the differing ones.
void launch_check_division(int64_t samples, uint64_t seed, unsigned long long* bad,
                           hipStream_t s);
// d0[slot] = |x0_p - x0_col[slot]| for every slot below the row's capacity;
// *bad = 1 if one lies outside the shared quotients' range, else 0
void launch_bond_d0(int32_t n, const int64_t* slice_off, const int32_t* cap, const int32_t* col,
                    const double4* x0, double* d0, unsigned int* bad, hipStream_t s);
// contact forces of the m candidates cand[0..m) (particles 0..m when cand
// is null); fc of every other particle is set to zero (fc_zeroed: the zone
// select already cleared it)
void launch_contact(int32_t n, const double4* x, const int32_t* body,
                    const int32_t* act, const GridView& g, const int32_t* cand,
                    int32_t m, double Rc, double kc, double vol, int all_pairs,
                    double4* fc, hipStream_t s, bool fc_zeroed = false);
void launch_reset_counters(pd_step_counters* c, hipStream_t s);

// Per-body boxes, already widened by the contact radius: particle p is a
// contact candidate if it lies in the box of a body other than its own.
struct ZoneBoxes {
  int32_t nbody;
  double lo[PD_MAX_ZONE_BODIES][3];
  double hi[PD_MAX_ZONE_BODIES][3];
};

// The contact sub-grid plan of a step, formed on the device from the body
// boxes: the zones, the componentwise bounds of all bodies (mn, mx: the
// grid origin), of all zones (ulo, uhi) and of the intersections of each
// zone with another body's box (glo, ghi; glo > ghi when none meet).
struct ContactPlan {
  ZoneBoxes z;
  double mn[3], mx[3], ulo[3], uhi[3], glo[3], ghi[3];
};

// pd_bonds.hip
// bounds of the particles of body `body_sel` (all particles if < 0)
// the bounding boxes of bodies 0 .. nbody - 1: out[b * 6 + k] (lo xyz, hi
// xyz; +inf / -inf for bodies without particles); part holds
// 1024 * PD_MAX_BODIES * 6
// plan non-null: the contact plan of the boxes too, zones widened by w
void launch_bbox_bodies(const double4* x, const int32_t* body, int32_t n, int32_t nbody,
                        double* part, double* out, hipStream_t s,
                        ContactPlan* plan = nullptr, double w = 0.0);
void launch_bbox(const double4* x, const int32_t* body, int32_t n, int32_t body_sel,
                 double* part, double* out6, hipStream_t s);
// stable (ascending) list of contact candidates; the count lands in
// tile_cnt[zone_tile_words(n) - 1]
// (fc non-null: every particle's contact force is cleared on the way)
void launch_zone_select(const double4* x, const int32_t* body, int32_t n,
                        const ZoneBoxes* z, uint32_t* tile_cnt, int32_t* out,
                        hipStream_t s, double4* fc = nullptr);

size_t zone_tile_words(int32_t n);
// keys of the cells (sub-grid origin cell c0, dims) of particles subset[i]
// (or i when subset is null); idx[i] receives the particle index
void launch_cell_keys(const double4* x, const int32_t* subset, int32_t n,
                      const double* mn, double cell, const int32_t* c0,
                      const int32_t* dims, uint32_t* keys, int32_t* idx,
                      hipStream_t s);
// Stable LSD radix sort of (keys, idx) by the low `bits` bits; returns which
// buffer pair (0: a, 1: b) holds the result.
int radix_sort(uint32_t* keys_a, int32_t* idx_a, uint32_t* keys_b, int32_t* idx_b,
               int32_t n, int32_t bits, uint32_t* hist, hipStream_t s);
size_t radix_hist_words(int32_t n);
void launch_cell_ranges(const uint32_t* keys, int32_t n, int32_t* start,
                        int32_t* end, int64_t ncell, hipStream_t s);
// Bounds of the one-workgroup build below
#define PD_SMALL_GRID_MAX_M 16384
#define PD_SMALL_GRID_MAX_CELLS 65536
// keys, stable sort and cell start / end of subset[0..m) in one launch
// (m <= PD_SMALL_GRID_MAX_M, ncell <= PD_SMALL_GRID_MAX_CELLS); returns which
// buffer pair (0: a, 1: b) holds the sorted keys and indices
int launch_small_grid(const double4* x, const int32_t* subset, int32_t m, const double* mn,
                      double cell, const int32_t* c0, const int32_t* dims, int32_t ncell,
                      int32_t bits, uint32_t* keys_a, int32_t* idx_a, uint32_t* keys_b,
                      int32_t* idx_b, int32_t* start, int32_t* end, hipStream_t s);
// INVISIBLE particles take no bonds and are no one's neighbour
void launch_bond_count(int32_t n, const double4* x, const int32_t* body, const int32_t* act,
                       const double* horizon, const GridView& g, int32_t* counts,
                       hipStream_t s);
void launch_slice_width(const int32_t* counts, int32_t n, int32_t* widths,
                        hipStream_t s);
void launch_bond_fill(int32_t n, const double4* x, const int32_t* body, const int32_t* act,
                      const double* horizon, const GridView& g, const int32_t* counts,
                      const int64_t* slice_off, int32_t* col, int32_t* top,
                      int32_t* cap, hipStream_t s);

}  // namespace pd
