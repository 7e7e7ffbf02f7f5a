// pd_api.hip -- the C ABI declared in include/pd_solver.h.
//
// A context owns every device buffer of one particle system and one HIP
// stream (default priority).  Inside a step the host waits only at the CG
// chunk boundaries (a pinned copy of the device-side CG state) and, when
// contact is on, once for the grid bounds.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "pd_internal.h"
#include "../../include/pd_solver.h"

static_assert(PD_LEARNED_PARAMS == PC_NPARAM, "learned parameter layout");

static thread_local std::string g_last_error;

struct pd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int32_t n = 0;
  int32_t nslices = 0;
  int32_t max_bonds = 528;
  int32_t nbody = 1;
  pd_material mat[PD_MAX_BODIES];
  std::string err;

  // particle state
  double4 *x = nullptr, *v = nullptr, *x0 = nullptr, *sol = nullptr, *f = nullptr,
          *b = nullptr, *fc = nullptr, *vbc = nullptr;
  double *s = nullptr, *s_new = nullptr, *s00 = nullptr;
  int32_t *act = nullptr, *body = nullptr, *bc = nullptr, *top = nullptr, *cap = nullptr;

  // bonds and matrix (sliced layout, see pd_device.h)
  int64_t* slice_off = nullptr;
  std::vector<int64_t> h_slice_off;
  std::vector<int32_t> h_cap;  // row capacities (initial bond counts)
  int64_t slots = 0;      // slot rows (each 64 lanes)
  int64_t slots_cap = 0;  // allocated slot rows
  int32_t* col = nullptr;
  double* val = nullptr;
  double* diag = nullptr;
  double* minv = nullptr;
  bool bonds_ready = false;
  int64_t live_bonds = 0;

  // mixed precision (allocated on the first assembly with precision 1)
  float* val32 = nullptr;
  float* diag32 = nullptr;
  float* minv32 = nullptr;
  int64_t slots32_cap = 0;
  float4 *r32 = nullptr, *z32 = nullptr, *p32 = nullptr, *y32 = nullptr, *d32 = nullptr;
  pd_mx_state* mx = nullptr;    // [nspec]
  pd_mx_state* h_mx = nullptr;  // pinned [nspec]
  int last_mx_iters = 0;
  int last_solve_prec = -1;
  int mat_prec = -1;  // precision of the last assembly (-1: none)
  int unroll[2] = {4, 4};  // SpMV slots per trip in use (FP64, mixed)
  int tuned[2] = {0, 0};   // pd_tune_spmv's choice (0: not tuned)

  // CG
  double4 *r = nullptr, *z = nullptr, *p = nullptr, *y = nullptr;
  double* part = nullptr;
  pd_cg_state* st = nullptr;  // [nspec]
  pd_step_counters* counters = nullptr;
  pd_mat_dev* dmat = nullptr;
  double* dhorizon = nullptr;
  int last_iters = 0;

  // bond-gap tracking
  double* gap_val = nullptr;
  int32_t* gap_q = nullptr;
  double* gpart_v = nullptr;
  int32_t* gpart_p = nullptr;
  int32_t* gpart_q = nullptr;
  double* gap_out = nullptr;
  int32_t* gap_pq = nullptr;

  // grid, sort
  uint32_t *keys_a = nullptr, *keys_b = nullptr;
  int32_t *idx_a = nullptr, *idx_b = nullptr;
  uint32_t* hist = nullptr;
  double* bbox_part = nullptr;
  double* bbox = nullptr;
  pd::ContactPlan* plan = nullptr;  // the step's contact plan (k_contact_plan)
  int32_t *cell_start = nullptr, *cell_end = nullptr;
  int64_t cell_cap = 0;
  int32_t* counts = nullptr;
  int32_t* widths = nullptr;
  // the last contact grid build: 0 none, 1 / 2 the one-workgroup build with
  // one / two radix passes, 3 the multi-launch build
  int last_grid = 0;
  uint32_t* zone_cnt = nullptr;  // contact candidate tile counts + total
  int32_t* cand = nullptr;       // contact candidates

  // rollback copies
  double4 *x_bak = nullptr, *v_bak = nullptr;
  int32_t* top_bak = nullptr;

  // independent specimens (one by default): row ranges and tile tables
  int32_t nspec = 1;
  int32_t state_cap = 0;  // entries allocated in st / h_st (and mx / h_mx)
  std::vector<int32_t> h_spec_first, h_spec_count;
  int32_t *tile_spec = nullptr, *spec_tile0 = nullptr, *spec_ntiles = nullptr;
  bool last_converged = false;  // every specimen of the last solve

  // pinned host mirrors
  pd_cg_state* h_st = nullptr;  // [nspec]
  pd_step_counters* h_cnt = nullptr;
  double* h_bbox = nullptr;
  double* h_gap = nullptr;
  int32_t* h_gap_pq = nullptr;
  uint32_t* h_zone_count = nullptr;
  pd::ContactPlan* h_plan = nullptr;

  std::vector<hipEvent_t> ev;  // event pool for timing

  // captured steps (graph = 1): assembly + solve start + K iterations, and
  // the position update; step scalars come from consts
  pd_step_consts* consts = nullptr;
  pd_step_consts* h_consts = nullptr;

  // learned preconditioner
  float2* bond_stats = nullptr;  // (mean, max) live-bond stretch per row
  float* pc_feat = nullptr;      // [n][PC_NIN] features (lazy)
  float* pc_p = nullptr;         // [PC_NPARAM]
  __bf16* pc_w = nullptr;        // [PC_NWEIGHT]
  bool pc_loaded = false;
  // the learned blocks M_p (FP64 planes, and their float copy for the mixed
  // solve), kept apart from the assembly's block-Jacobi inverse so that a
  // fallback continues with block Jacobi at once (allocated on first use)
  double* pc_minv = nullptr;
  float* pc_minv32 = nullptr;
  bool last_learned = false;  // the last solve applied the learned blocks
  int bj_iters = 0;  // iterations of the last block-Jacobi solve
  int pc_since_probe = 0;  // precond-2 steps since bj_iters was refreshed
  bool pc_fell_back = false;  // the last learned solve fell back to block Jacobi
  hipGraphExec_t g_solve = nullptr, g_update = nullptr;
  std::string graph_key;
  int graph_k = 0;

  // matrix-free operator: set by the last assembly
  int mf = 0;  // 1: the last assembly stored no bond blocks (matrix-free solve)
  // reference bond lengths |x0_p - x0_q| per slot, moved with the columns by
  // the swap-removes; recomputed when bonds or reference positions change
  double* d0 = nullptr;
  bool d0_valid = false;
  unsigned int* d0_bad = nullptr;  // some d0 outside the shared quotients' range
  // the assembly record, per particle (x, y, z, s, xs.x, xs.y, xs.z, 0)
  double* arec = nullptr;
  double asm_dt2 = 0.0;                      // its dt^2 scale
  const pd_step_consts* asm_consts = nullptr;  // captured step: dt2 read here
  // per particle the position and the CG direction side by side (6 doubles,
  // the matrix-free SpMV's gather record), allocated on first use
  double* rec = nullptr;
  // true from an assembly until positions change (update, upload, restore,
  // new bonds): the assembly's range flag and record describe the positions
  bool asm_fresh = false;
  // a stored-block FP64 step's assembly left the blocks to the solve's first
  // product (launch_cg_spmv_store); cleared when that product or
  // ensure_blocks has written them
  bool blocks_deferred = false;
  // the same for the float blocks of a mixed assembly (launch_mx_spmv_store)
  bool blocks32_deferred = false;
  // matrix_free = 2 (pd_step): measured times of the two operators, ms per
  // CG product (0: not measured yet); reset when the bonds change
  double op_t_mf = 0.0;     // recomputing product
  double op_t_store = 0.0;  // the storing first product of a stored-block step
  double op_t_st = 0.0;     // a streaming product of stored blocks
};

struct pd_trainer {
  int device = 0;
  hipStream_t stream = nullptr;
  float* samples = nullptr;  // [cap][PC_SAMPLE]
  int64_t cap = 0;
  int64_t count = 0;
  float *p = nullptr, *m = nullptr, *v = nullptr;  // parameters, Adam moments
  __bf16* w = nullptr;
  float* partial = nullptr;
  int32_t partial_waves = 0;
  float* loss = nullptr;
  int32_t loss_cap = 0;
  int32_t steps_done = 0;
  bool normed = false;
  std::string err;
};

namespace {

int fail(pd_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  g_last_error = msg;
  return code;
}

#define HIPCHK(ctx, expr)                                                                \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return fail((ctx), PD_ERR_HIP,                                                     \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                    \
  } while (0)

#define CHECK_LAUNCH(ctx, what)                                                          \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return fail((ctx), PD_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e_));  \
  } while (0)

template <typename T>
int dalloc(pd_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess)
    return fail(c, PD_ERR_HIP, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                   " B): " + hipGetErrorString(e));
  e = hipMemsetAsync(*p, 0, count * sizeof(T), c->stream);
  if (e != hipSuccess) return fail(c, PD_ERR_HIP, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  return PD_OK;
}

template <typename T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

struct FieldInfo {
  void** dptr;
  size_t elem;  // bytes per particle
  bool vec4;
  bool writable;
};

FieldInfo field_info(pd_ctx* c, int32_t f) {
  switch (f) {
    case PD_FIELD_POSITION: return {(void**)&c->x, 32, true, true};
    case PD_FIELD_VELOCITY: return {(void**)&c->v, 32, true, true};
    case PD_FIELD_ORIGINAL_POSITION: return {(void**)&c->x0, 32, true, true};
    case PD_FIELD_STRETCH: return {(void**)&c->s, 8, false, true};
    case PD_FIELD_ORIGINAL_STRETCH: return {(void**)&c->s00, 8, false, true};
    case PD_FIELD_ACTIVITY: return {(void**)&c->act, 4, false, true};
    case PD_FIELD_BODY: return {(void**)&c->body, 4, false, true};
    case PD_FIELD_BC_FLAG: return {(void**)&c->bc, 4, false, true};
    case PD_FIELD_BC_VELOCITY: return {(void**)&c->vbc, 32, true, true};
    case PD_FIELD_SOLUTION: return {(void**)&c->sol, 32, true, true};
    case PD_FIELD_FORCE: return {(void**)&c->f, 32, true, false};
    case PD_FIELD_RHS: return {(void**)&c->b, 32, true, false};
    case PD_FIELD_NEW_STRETCH: return {(void**)&c->s_new, 8, false, false};
    case PD_FIELD_BOND_TOP: return {(void**)&c->top, 4, false, true};
    case PD_FIELD_CONTACT_FORCE: return {(void**)&c->fc, 32, true, false};
    default: return {nullptr, 0, false, false};
  }
}

int ensure_events(pd_ctx* c, size_t count) {
  while (c->ev.size() < count) {
    hipEvent_t e;
    HIPCHK(c, hipEventCreate(&e));
    c->ev.push_back(e);
  }
  return PD_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

int check_params(pd_ctx* c, const pd_step_params* prm) {
  if (!prm) return fail(c, PD_ERR_ARG, "null step parameters");
  if (!(prm->dt > 0.0)) return fail(c, PD_ERR_ARG, "dt must be > 0");
  if (prm->integrator != 0 && prm->integrator != 1)
    return fail(c, PD_ERR_ARG, "integrator must be 0 (implicit Euler) or 1 (Newmark-beta)");
  if (prm->precond < 0 || prm->precond > 2)
    return fail(c, PD_ERR_ARG, "precond must be 0 (identity), 1 (block Jacobi) or 2 (learned)");
  if (prm->precond == 2 && !c->pc_loaded)
    return fail(c, PD_ERR_STATE, "precond 2 needs a network (pd_learned_set)");
  if (prm->precond == 2 && prm->graph)
    return fail(c, PD_ERR_ARG, "captured steps (graph = 1) support precond 0 and 1");
  if (prm->contact_radius > 0.0 && !(prm->contact_cell >= prm->contact_radius))
    return fail(c, PD_ERR_ARG, "contact_cell must be >= contact_radius");
  if (prm->contact_radius > 0.0 && !(prm->contact_volume != 0.0))
    return fail(c, PD_ERR_ARG, "contact_volume must be non-zero");
  if (prm->contact_mode != 0 && prm->contact_mode != 1)
    return fail(c, PD_ERR_ARG, "contact_mode must be 0 (between bodies) or 1 (all pairs)");
  if (prm->spmv_unroll != 0 && prm->spmv_unroll != 2 && prm->spmv_unroll != 4 &&
      prm->spmv_unroll != 8)
    return fail(c, PD_ERR_ARG, "spmv_unroll must be 0 (tuned / default), 2, 4 or 8");
  for (int k = 0; k < 2; ++k)
    c->unroll[k] = prm->spmv_unroll ? prm->spmv_unroll : (c->tuned[k] ? c->tuned[k] : 4);
  if (prm->precision != 0 && prm->precision != 1)
    return fail(c, PD_ERR_ARG, "precision must be 0 (FP64) or 1 (float inner solve + FP64 refinement)");
  if (prm->matrix_free < 0 || prm->matrix_free > 2)
    return fail(c, PD_ERR_ARG,
                "matrix_free must be 0 (stored blocks), 1 (recomputed blocks) or 2 (per step)");
  if (prm->matrix_free == 1 && prm->precision != 0)
    return fail(c, PD_ERR_ARG, "matrix_free needs precision 0 (the float solve streams float blocks)");
  if (prm->precision == 1 && (prm->inner_rtol < 0.0 || prm->max_refine < 0 || prm->inner_max_iter < 0))
    return fail(c, PD_ERR_ARG, "inner_rtol, max_refine and inner_max_iter must be >= 0");
  if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built (pd_build_bonds / pd_set_bonds)");
  return PD_OK;
}

// Buffers of the mixed-precision path, sized for the current slot capacity.
int ensure_mixed(pd_ctx* c) {
  const size_t n = (size_t)c->n;
  const size_t npad = (size_t)c->nslices * 64;
  if (!c->mx) {
    int e;
    if ((e = dalloc(c, &c->diag32, npad * 6)) || (e = dalloc(c, &c->minv32, npad * 6)) ||
        (e = dalloc(c, &c->r32, n)) || (e = dalloc(c, &c->z32, n)) ||
        (e = dalloc(c, &c->p32, n)) || (e = dalloc(c, &c->y32, n)) ||
        (e = dalloc(c, &c->d32, n)) || (e = dalloc(c, &c->mx, (size_t)c->state_cap)))
      return e;
    if (hipHostMalloc((void**)&c->h_mx, sizeof(pd_mx_state) * c->state_cap,
                      hipHostMallocDefault) != hipSuccess)
      return fail(c, PD_ERR_HIP, "hipHostMalloc failed");
    memset(c->h_mx, 0, sizeof(pd_mx_state) * c->state_cap);
  }
  if (c->slots32_cap < c->slots_cap) {
    dfree(c->val32);
    if (int e = dalloc(c, &c->val32, (size_t)c->slots_cap * 64 * 6)) return e;
    c->slots32_cap = c->slots_cap;
  }
  return PD_OK;
}

// Mixed-precision limits with the documented defaults for zero fields.
struct MixedLimits {
  double inner_rtol;
  int32_t max_refine;
  int32_t inner_max;
};
MixedLimits mixed_limits(const pd_step_params* prm) {
  MixedLimits m;
  m.inner_rtol = prm->inner_rtol > 0.0 ? prm->inner_rtol : 1e-5;
  m.max_refine = prm->max_refine > 0 ? prm->max_refine : 20;
  m.inner_max = prm->inner_max_iter > 0 ? prm->inner_max_iter : prm->max_iter;
  return m;
}

// Bounds of the particles of body `sel` (all when sel < 0), read back.
int read_bbox(pd_ctx* c, int32_t sel, double* out6) {
  pd::launch_bbox(c->x, c->body, c->n, sel, c->bbox_part, c->bbox, c->stream);
  CHECK_LAUNCH(c, "bbox");
  HIPCHK(c, hipMemcpyAsync(c->h_bbox, c->bbox, 6 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  memcpy(out6, c->h_bbox, 6 * sizeof(double));
  return PD_OK;
}

// bounding boxes of bodies 0 .. nbody - 1, one pass and one wait
int read_bboxes(pd_ctx* c, double (*out)[6]) {
  pd::launch_bbox_bodies(c->x, c->body, c->n, c->nbody, c->bbox_part, c->bbox, c->stream);
  CHECK_LAUNCH(c, "bbox");
  HIPCHK(c, hipMemcpyAsync(c->h_bbox, c->bbox, 6 * PD_MAX_BODIES * sizeof(double),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int b = 0; b < c->nbody; ++b) memcpy(out[b], c->h_bbox + 6 * b, 6 * sizeof(double));
  return PD_OK;
}

// Uniform grid with origin mn and edge `cell`: the sub-grid of cells
// c0 .. c0 + dims - 1 over the m particles subset[0..m) (all n particles
// when subset is null), sorted by cell with a stable radix sort so each
// cell lists its particles in ascending index, plus cell start / end.
int grid_over(pd_ctx* c, double cell, const double* mn, const int32_t* c0, const int32_t* dims,
              const int32_t* subset, int32_t m, pd::GridView* g) {
  int64_t ncell = 1;
  for (int k = 0; k < 3; ++k) {
    g->mn[k] = mn[k];
    g->c0[k] = c0[k];
    g->dims[k] = dims[k];
    ncell *= dims[k];
  }
  if (ncell > (int64_t)1 << 31) return fail(c, PD_ERR_ARG, "grid has more than 2^31 cells");
  g->cell = cell;
  if (ncell > c->cell_cap) {
    dfree(c->cell_start);
    dfree(c->cell_end);
    const int64_t cap = std::max<int64_t>(ncell, c->cell_cap * 3 / 2);
    if (int e = dalloc(c, &c->cell_start, (size_t)cap)) return e;
    if (int e = dalloc(c, &c->cell_end, (size_t)cap)) return e;
    c->cell_cap = cap;
  }
  int bits = 1;
  while (bits < 32 && ((int64_t)1 << bits) < ncell) ++bits;
  // a candidate subset of contact size: the whole build in one launch
  const bool small = subset && m > 0 && m <= PD_SMALL_GRID_MAX_M &&
                     ncell <= PD_SMALL_GRID_MAX_CELLS;
  c->last_grid = small ? (bits <= 8 ? 1 : 2) : 3;
  int where;
  if (small) {
    where = pd::launch_small_grid(c->x, subset, m, mn, cell, c0, dims, (int32_t)ncell, bits,
                                  c->keys_a, c->idx_a, c->keys_b, c->idx_b, c->cell_start,
                                  c->cell_end, c->stream);
    CHECK_LAUNCH(c, "small grid");
  } else {
    pd::launch_cell_keys(c->x, subset, m, mn, cell, c0, dims, c->keys_a, c->idx_a, c->stream);
    CHECK_LAUNCH(c, "cell keys");
    where = pd::radix_sort(c->keys_a, c->idx_a, c->keys_b, c->idx_b, m, bits, c->hist,
                           c->stream);
    CHECK_LAUNCH(c, "radix sort");
  }
  const uint32_t* keys = where ? c->keys_b : c->keys_a;
  const int32_t* idx = where ? c->idx_b : c->idx_a;
  if (!small) {
    pd::launch_cell_ranges(keys, m, c->cell_start, c->cell_end, ncell, c->stream);
    CHECK_LAUNCH(c, "cell ranges");
  }
  g->sorted_key = keys;
  g->sorted_idx = idx;
  g->cell_start = c->cell_start;
  g->cell_end = c->cell_end;
  return PD_OK;
}

int cell_index(double x, double mn, double cell) { return (int)floor((x - mn) / cell); }

// Grid over all particles (bond build): origin at the componentwise minimum.
int build_grid(pd_ctx* c, double cell, pd::GridView* g) {
  double bb[6];
  if (int e = read_bbox(c, -1, bb)) return e;
  int32_t c0[3] = {0, 0, 0}, dims[3];
  for (int k = 0; k < 3; ++k) {
    if (!(isfinite(bb[k]) && isfinite(bb[3 + k])))
      return fail(c, PD_ERR_ARG, "non-finite particle position");
    const double span = floor((bb[3 + k] - bb[k]) / cell) + 1.0;
    if (span > 2.0e9) return fail(c, PD_ERR_ARG, "grid too large for the cell size");
    dims[k] = (int32_t)span;
  }
  return grid_over(c, cell, bb, c0, dims, nullptr, c->n, g);
}
int ensure_slots(pd_ctx* c, int64_t slots) {
  if (slots <= c->slots_cap) return PD_OK;
  dfree(c->col);
  dfree(c->val);
  dfree(c->d0);
  c->d0_valid = false;
  c->blocks_deferred = false;  // new bond lists: no matrix to keep
  c->blocks32_deferred = false;
  c->asm_fresh = false;
  if (int e = dalloc(c, &c->col, (size_t)slots * 64)) return e;
  if (int e = dalloc(c, &c->val, (size_t)slots * 64 * 6)) return e;
  if (int e = dalloc(c, &c->d0, (size_t)slots * 64)) return e;
  c->slots_cap = slots;
  return PD_OK;
}

int set_slice_offsets(pd_ctx* c, const std::vector<int32_t>& widths) {
  c->h_slice_off.assign(c->nslices + 1, 0);
  for (int s = 0; s < c->nslices; ++s) c->h_slice_off[s + 1] = c->h_slice_off[s] + widths[s];
  c->slots = c->h_slice_off[c->nslices];
  if (int e = ensure_slots(c, c->slots)) return e;
  HIPCHK(c, hipMemcpyAsync(c->slice_off, c->h_slice_off.data(),
                           sizeof(int64_t) * (c->nslices + 1), hipMemcpyHostToDevice, c->stream));
  return PD_OK;
}

pd::SpmvView spmv_view(pd_ctx* c) {
  pd::SpmvView A;
  A.n = c->n;
  A.slice_off = c->slice_off;
  A.top = c->top;
  A.col = c->col;
  A.val = c->val;
  A.diag = c->diag;
  A.unroll = c->unroll[0];
  A.mf = c->mf;
  A.pos = c->x;
  A.body = c->body;
  A.mat = c->dmat;
  A.dt2 = c->asm_dt2;
  A.consts = c->asm_consts;
  A.d0 = c->d0;
  A.range_bad = c->asm_fresh ? (const unsigned int*)((const char*)c->counters +
                                                     offsetof(pd_step_counters, range_bad))
                             : nullptr;
  A.rec = c->asm_fresh && c->mf ? c->rec : nullptr;
  return A;
}

pd::Batch batch_view(pd_ctx* c) {
  pd::Batch b;
  b.nspec = c->nspec;
  const bool many = c->tile_spec != nullptr;
  b.tile_spec = many ? c->tile_spec : nullptr;
  b.spec_tile0 = many ? c->spec_tile0 : nullptr;
  b.spec_ntiles = many ? c->spec_ntiles : nullptr;
  return b;
}

pd::MxArgs mx_args(pd_ctx* c, int precond) {
  pd::MxArgs a;
  a.n = c->n;
  a.batch = batch_view(c);
  a.slice_off = c->slice_off;
  a.top = c->top;
  a.col = c->col;
  a.val32 = c->val32;
  a.diag = c->diag;
  a.diag32 = c->diag32;
  a.minv32 = precond == 2 ? c->pc_minv32 : c->minv32;
  a.bc = c->bc;
  a.b = c->b;
  a.x = c->sol;
  a.r = c->r32;
  a.z = c->z32;
  a.p = c->p32;
  a.y = c->y32;
  a.d = c->d32;
  a.part = c->part;
  a.st = c->mx;
  a.precond = precond;
  a.unroll = c->unroll[1];
  a.A64 = spmv_view(c);
  // the mixed assembly wrote the record's positions (the residual gathers
  // positions and x from it)
  a.A64.rec = c->asm_fresh && c->mat_prec == 1 && c->rec ? c->rec : nullptr;
  return a;
}

pd::CgArgs cg_args(pd_ctx* c, int precond) {
  pd::CgArgs a;
  a.A = spmv_view(c);
  a.batch = batch_view(c);
  a.minv = precond == 2 ? c->pc_minv : c->minv;
  a.bc = c->bc;
  a.b = c->b;
  a.x = c->sol;
  a.r = c->r;
  a.z = c->z;
  a.p = c->p;
  a.y = c->y;
  a.part = c->part;
  a.st = c->st;
  a.precond = precond;
  a.vals = c->val;
  a.rec = c->rec;
  return a;
}

// Contact: only particles within the contact radius of another body's
// bounding box can touch it, so the per-step grid is built over those
// candidates alone.  The grid keeps the origin (minimum over all particles)
// and cell edges of the full grid, and candidates stay in ascending order,
// so every particle sums its contacts in the same order as on the full grid.
int enqueue_contact(pd_ctx* c, const pd_step_params* prm) {
  c->last_grid = 0;
  const int nb = c->nbody;
  const double cell = prm->contact_cell;
  if (prm->contact_mode == 1) {
    // every pair: the grid over all particles
    double bb[PD_MAX_BODIES][6];
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (int e = read_bboxes(c, bb)) return e;
    for (int b = 0; b < nb; ++b) {
      for (int k = 0; k < 3; ++k) {
        mn[k] = std::min(mn[k], bb[b][k]);
        mx[k] = std::max(mx[k], bb[b][3 + k]);
      }
    }
    for (int k = 0; k < 3; ++k)
      if (!(isfinite(mn[k]) && isfinite(mx[k])))
        return fail(c, PD_ERR_ARG, "non-finite particle position");
    pd::GridView g = {};
    int32_t c0[3] = {0, 0, 0}, dims[3];
    for (int k = 0; k < 3; ++k) {
      const double span = floor((mx[k] - mn[k]) / cell) + 1.0;
      if (span > 2.0e9) return fail(c, PD_ERR_ARG, "contact grid too large for contact_cell");
      dims[k] = (int32_t)span;
    }
    if (int e = grid_over(c, cell, mn, c0, dims, nullptr, c->n, &g)) return e;
    pd::launch_contact(c->n, c->x, c->body, c->act, g, nullptr, c->n, prm->contact_radius,
                       prm->contact_stiffness, prm->contact_volume, 1, c->fc, c->stream);
    CHECK_LAUNCH(c, "contact");
    return PD_OK;
  }
  // The body boxes, the zones (boxes widened by the contact radius, a small
  // widening keeping pairs at a distance rounding to Rc inside) and the
  // sub-grid box are formed on the device and read back with the candidate
  // count, in one host wait.  A candidate of body o lies in o's own box and
  // in the zone of another body b, so every candidate's cell lies in the
  // box over those intersections; the sub-grid spans just that box.  Its
  // cells keep their full-grid coordinates and the cells it leaves out hold
  // no candidate, so each contact sum runs over the same members in the
  // same order.
  const double w = prm->contact_radius * (1.0 + 1e-6);
  pd::launch_bbox_bodies(c->x, c->body, c->n, c->nbody, c->bbox_part, c->bbox, c->stream,
                         c->plan, w);
  const size_t tw = pd::zone_tile_words(c->n);
  pd::launch_zone_select(c->x, c->body, c->n, &c->plan->z, c->zone_cnt, c->cand, c->stream,
                         c->fc);
  CHECK_LAUNCH(c, "contact candidates");
  HIPCHK(c, hipMemcpyAsync(c->h_plan, c->plan, sizeof(pd::ContactPlan), hipMemcpyDeviceToHost,
                           c->stream));
  uint32_t m = 0;
  if (c->n > 0)
    HIPCHK(c, hipMemcpyAsync(c->h_zone_count, c->zone_cnt + tw - 1, sizeof(uint32_t),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const pd::ContactPlan& P = *c->h_plan;
  for (int k = 0; k < 3; ++k)
    if (!(isfinite(P.mn[k]) && isfinite(P.mx[k])))
      return fail(c, PD_ERR_ARG, "non-finite particle position");
  if (c->n > 0) m = *c->h_zone_count;
  pd::GridView g = {};
  int32_t c0[3] = {0, 0, 0}, dims[3] = {1, 1, 1};
  if (m > 0) {
    for (int k = 0; k < 3; ++k) {
      const bool tight = P.glo[k] <= P.ghi[k];  // always, when a candidate exists
      const double lo = tight ? P.glo[k] : std::max(P.ulo[k], P.mn[k]);
      const double hi = tight ? P.ghi[k] : std::min(P.uhi[k], P.mx[k]);
      c0[k] = cell_index(lo, P.mn[k], cell);
      dims[k] = cell_index(hi, P.mn[k], cell) - c0[k] + 1;
      if (dims[k] < 1) dims[k] = 1;
    }
    if (int e = grid_over(c, cell, P.mn, c0, dims, c->cand, (int32_t)m, &g)) return e;
  }
  pd::launch_contact(c->n, c->x, c->body, c->act, g, c->cand, (int32_t)m, prm->contact_radius,
                     prm->contact_stiffness, prm->contact_volume, 0, c->fc, c->stream, true);
  CHECK_LAUNCH(c, "contact");
  return PD_OK;
}

int ensure_rec(pd_ctx* c) {
  if (!c->rec)
    if (int e = dalloc(c, &c->rec, (size_t)c->n * 6)) return e;
  return PD_OK;
}

// FP64 stored blocks: the solve's first product writes them (the CG
// product when the start is fused into the assembly, else the start's
// y = A x0), at a lower cost than the assembly's stores add
bool defer_blocks(const pd_step_params* prm) {
  return !prm->matrix_free && prm->precision == 0;
}

// Deferred blocks written from the assembly's positions (for callers that
// read the stored blocks without a solve having run).
int ensure_blocks(pd_ctx* c) {
  if (c->blocks_deferred) {
    pd::launch_write_blocks(spmv_view(c), c->val, c->stream);
    CHECK_LAUNCH(c, "write blocks");
    c->blocks_deferred = false;
  }
  if (c->blocks32_deferred) {
    pd::launch_write_blocks32(spmv_view(c), c->val32, c->stream);
    CHECK_LAUNCH(c, "write float blocks");
    c->blocks32_deferred = false;
  }
  return PD_OK;
}

// fuse: also produce the solve's first residual and start, from the blocks
// and the block-Jacobi inverse the pass has in registers (the next solve is
// run with fused = true).  With the learned preconditioner (precond 2) the
// start is the block-Jacobi one; run_solve_pc replaces z = M r and its
// partial once the learned blocks exist (launch_pc_infer_start), or keeps it for a
// block-Jacobi probe.
int enqueue_assemble(pd_ctx* c, const pd_step_params* prm,
                     const pd_step_consts* consts = nullptr, bool fuse = false) {
  if (prm->precision == 1)
    if (int e = ensure_mixed(c)) return e;
  pd::launch_reset_counters(c->counters, c->stream);
  const double dt2 = prm->integrator == 1 ? (prm->dt * prm->dt) / 4.0 : prm->dt * prm->dt;
  if (!c->d0_valid) {  // the assembly and the matrix-free SpMV read them
    pd::launch_bond_d0(c->n, c->slice_off, c->cap, c->col, c->x0, c->d0, c->d0_bad, c->stream);
    CHECK_LAUNCH(c, "reference bond lengths");
    c->d0_valid = true;
  }
  const bool defer = defer_blocks(prm);
  // the record: positions and p for the matrix-free FP64 products,
  // positions and x for the mixed solve's FP64 residual
  const bool rec = prm->matrix_free || defer || prm->precision == 1;
  if (rec)
    if (int e = ensure_rec(c)) return e;
  {  // the assembly record, the Dirichlet start values (fuse), the range flag
    pd::AsmPrepArgs pa;
    pa.n = c->n;
    pa.fuse = fuse ? 1 : 0;
    pa.integrator = prm->integrator;
    pa.c_bad = 0;
    auto in_range = [](double v) { return v >= 0x1p-300 && v <= 0x1p300; };
    for (int b = 0; b < c->nbody; ++b)
      if (!in_range(c->mat[b].c_bond) || !in_range(c->mat[b].horizon) ||
          !in_range(c->mat[b].radius))
        pa.c_bad = 1;
    pa.dt = prm->dt;
    pa.consts = consts;
    pa.x = c->x;
    pa.s = c->s;
    pa.sol = c->sol;
    pa.bc = c->bc;
    pa.act = c->act;
    pa.vbc = c->vbc;
    pa.d0_bad = c->d0_bad;
    pa.arec = c->arec;
    pa.rec = rec ? c->rec : nullptr;
    pa.counters = c->counters;
    pd::launch_asm_prep(pa, c->stream);
    CHECK_LAUNCH(c, "assembly record");
  }
  pd::AsmArgs a;
  a.d0 = c->d0;
  a.arec = c->arec;
  a.rec = prm->matrix_free || defer ? c->rec : nullptr;
  c->mf = prm->matrix_free ? 1 : 0;
  c->blocks_deferred = defer;
  c->blocks32_deferred = prm->precision == 1;  // launch_mx_spmv_store writes them
  c->asm_dt2 = dt2;
  c->asm_consts = consts;
  a.fuse = fuse ? 1 : 0;
  a.pc = prm->precond != 0 ? 1 : 0;  // precond 2: the block-Jacobi start first
  a.xs = c->sol;
  a.r64 = c->r;
  a.p64 = prm->matrix_free ? nullptr : c->p;  // matrix-free: p lives in the record alone
  a.r32 = c->r32;
  a.p32 = c->p32;
  a.d32 = c->d32;
  a.part = c->part;
  a.mixed = prm->precision == 1 ? 1 : 0;
  a.diag32 = c->diag32;
  a.minv32 = c->minv32;
  a.n = c->n;
  a.slice_off = c->slice_off;
  a.top = c->top;
  a.col = c->col;
  a.val = c->val;
  a.diag = c->diag;
  a.minv = c->minv;
  a.x = c->x;
  a.v = c->v;
  a.fc = c->fc;
  a.vbc = c->vbc;
  a.f = c->f;
  a.b = c->b;
  a.sol = c->sol;
  a.s = c->s;
  a.s_new = c->s_new;
  a.s00 = c->s00;
  a.act = c->act;
  a.body = c->body;
  a.bc = c->bc;
  a.mat = c->dmat;
  a.dt = prm->dt;
  a.dt2 = dt2;
  a.clock = prm->clock_scale;
  a.gravity = prm->gravity;
  a.integrator = prm->integrator;
  a.contact = prm->contact_radius > 0.0 ? 1 : 0;
  a.track_gap = prm->track_gap ? 1 : 0;
  a.gap_val = c->gap_val;
  a.gap_q = c->gap_q;
  a.counters = c->counters;
  a.consts = consts;
  a.bond_stats = c->bond_stats;
  pd::launch_assemble(a, c->stream);
  CHECK_LAUNCH(c, "assemble");
  c->mat_prec = prm->precision;
  c->asm_fresh = true;
  if (prm->track_gap) {
    pd::launch_gap_reduce(c->gap_val, c->gap_q, c->n, c->gpart_v, c->gpart_p, c->gpart_q,
                          c->gap_out, c->gap_pq, c->stream);
    CHECK_LAUNCH(c, "gap reduce");
  }
  HIPCHK(c, hipMemcpyAsync(c->h_cnt, c->counters, sizeof(pd_step_counters),
                           hipMemcpyDeviceToHost, c->stream));
  if (prm->track_gap) {
    HIPCHK(c, hipMemcpyAsync(c->h_gap, c->gap_out, sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->h_gap_pq, c->gap_pq, 2 * sizeof(int32_t), hipMemcpyDeviceToHost,
                             c->stream));
  }
  return PD_OK;
}

void read_assemble_stats(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st) {
  c->live_bonds = (int64_t)c->h_cnt->live;
  if (!st) return;
  st->broken_bonds = (int64_t)c->h_cnt->broken;
  st->live_bonds = (int64_t)c->h_cnt->live;
  if (prm->track_gap) {
    st->min_gap = *c->h_gap;
    st->gap_p = c->h_gap_pq[0];
    st->gap_q = c->h_gap_pq[1];
  } else {
    st->min_gap = INFINITY;
    st->gap_p = -1;
    st->gap_q = -1;
  }
}

// CG on the device; the host waits once per chunk of iterations.
int run_solve_mixed(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st, int resumed,
                    bool fused);

// resumed > 0: a captured step already ran the start and `resumed`
// iterations; the device state is in h_st / h_mx.  No per-launch timing then.
// fused: the assembly produced the mixed solve's first residual partials.
int run_solve(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st, int resumed = 0,
              bool fused = false) {
  if (c->mat_prec != prm->precision)
    return fail(c, PD_ERR_STATE, "the matrix was not assembled with this precision");
  if (prm->precision == 1) return run_solve_mixed(c, prm, st, resumed, fused);
  const pd::CgArgs a = cg_args(c, prm->precond);
  const int ns = c->nspec;
  if (!resumed) {
    if (fused) {  // the fold sets the solver state up too
      pd::launch_cg_init_fold(a, prm->rtol, prm->max_iter, c->stream);
    } else if (c->blocks_deferred) {  // the start's product writes the blocks
      pd::launch_cg_setup(c->st, ns, prm->rtol, prm->max_iter, c->stream);
      pd::launch_cg_init_store(a, c->stream);
      c->blocks_deferred = false;
    } else {
      pd::launch_cg_setup(c->st, ns, prm->rtol, prm->max_iter, c->stream);
      pd::launch_cg_init(a, c->stream);
    }
    CHECK_LAUNCH(c, "cg init");
  }
  int launched = resumed;
  int syncs = 0;
  int store_at = -1;  // the launch that also wrote the blocks
  // launch as many iterations as the previous solve needed (steps in a run
  // mostly repeat their count, and a launch after convergence returns at
  // once, cheaper than a host round trip), then single iterations, then
  // doubling chunks
  int chunk = resumed ? 1 : std::max(1, c->last_iters);
  int next = 1;
  int singles = 2;
  const bool timing = prm->timing != 0 && !resumed;
  bool all_done = false;
  if (resumed) {
    all_done = true;
    for (int k = 0; k < ns; ++k) all_done = all_done && c->h_st[k].done;
  }
  while (!all_done) {
    if (timing)
      if (int e = ensure_events(c, 2 * (size_t)(launched + chunk) + 8)) return e;
    for (int i = 0; i < chunk; ++i) {
      if (timing) HIPCHK(c, hipEventRecord(c->ev[8 + 2 * (launched + i)], c->stream));
      if (c->blocks_deferred) {  // the first product writes the blocks
        pd::launch_cg_spmv_store(a, c->stream);
        c->blocks_deferred = false;
        store_at = launched + i;
      } else {
        pd::launch_cg_spmv(a, c->stream);
      }
      if (timing) HIPCHK(c, hipEventRecord(c->ev[8 + 2 * (launched + i) + 1], c->stream));
      pd::launch_cg_update(a, launched + i, c->stream);
      pd::launch_cg_dir(a, launched + i, c->stream);
    }
    CHECK_LAUNCH(c, "cg iteration");
    launched += chunk;
    HIPCHK(c, hipMemcpyAsync(c->h_st, c->st, sizeof(pd_cg_state) * ns, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ++syncs;
    all_done = true;
    for (int k = 0; k < ns; ++k) all_done = all_done && c->h_st[k].done;
    if (all_done) break;
    if (launched > prm->max_iter + 1)
      return fail(c, PD_ERR_HIP, "CG state did not terminate on the device");
    if (singles > 0) {
      --singles;
      chunk = 1;
    } else {
      next = std::min(next * 2, 32);
      chunk = next;
    }
  }
  // the batch summary: most iterations, all converged, worst relative residual
  int it = 0, worst = 0;
  bool conv = true;
  for (int k = 0; k < ns; ++k) {
    const pd_cg_state& h = c->h_st[k];
    it = std::max(it, h.iter);
    conv = conv && h.converged;
    const pd_cg_state& w = c->h_st[worst];
    if (h.rnorm * w.bnorm > w.rnorm * h.bnorm) worst = k;
  }
  c->last_iters = it;
  c->last_converged = conv;
  c->last_solve_prec = 0;
  if (st) {
    st->cg_iterations = it;
    st->converged = conv ? 1 : 0;
    st->residual_norm = c->h_st[worst].rnorm;
    st->b_norm = c->h_st[worst].bnorm;
    st->refinements = 0;
    st->host_syncs += syncs;
    if (timing) {
      double tot = 0.0;
      int cnt = 0;
      const int ran = std::min(it, launched);
      for (int k = 0; k < ran; ++k) {
        const double ms = elapsed(c->ev[8 + 2 * k], c->ev[8 + 2 * k + 1]);
        if (k == store_at) {
          st->ms_spmv_store = ms;
        } else {
          tot += ms;
          ++cnt;
        }
      }
      st->ms_spmv_total = tot;
      st->spmv_launches = cnt;
    }
  }
  return PD_OK;
}

// Mixed-precision solve on the device.  Each chunk enqueues `chunk` inner
// iterations, then the correction, the FP64 residual and its test; every
// kernel returns at once unless the device-side phase asks for it, so the
// host only predicts how many inner iterations to enqueue before it waits.
int run_solve_mixed(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st, int resumed,
                    bool fused) {
  const pd::MxArgs a = mx_args(c, prm->precond);
  const MixedLimits lim = mixed_limits(prm);
  const bool timing = prm->timing != 0 && !resumed;
  const int ns = c->nspec;
  // event pairs: [0] first residual, then per chunk `chunk` inner SpMVs and
  // one residual
  std::vector<int> spmv_ev, resid_ev;
  size_t used = 8;
  auto pair = [&](std::vector<int>& v) -> int {
    if (int e = ensure_events(c, used + 2)) return e;
    v.push_back((int)used);
    used += 2;
    return PD_OK;
  };
  if (!resumed) {
    pd::launch_mx_setup(c->mx, ns, prm->rtol, prm->max_iter, lim.inner_rtol, lim.inner_max,
                        lim.max_refine, c->stream);
    if (fused) {
      pd::launch_mx_resid_fold(a, c->stream);
    } else {
      pd::launch_mx_rec_x(a, c->stream);
      if (timing) {
        if (int e = pair(resid_ev)) return e;
        HIPCHK(c, hipEventRecord(c->ev[resid_ev.back()], c->stream));
      }
      pd::launch_mx_resid(a, c->stream);
      if (timing) HIPCHK(c, hipEventRecord(c->ev[resid_ev.back() + 1], c->stream));
    }
    CHECK_LAUNCH(c, "mixed residual");
  }
  int chunk = resumed ? 1 : std::max(1, c->last_mx_iters);
  int next = 1;
  int singles = 2;
  int syncs = 0;
  std::vector<int> iter_seen(ns, 0), resid_seen(ns, 1);
  bool all_done = false;
  if (resumed) {
    all_done = true;
    for (int k = 0; k < ns; ++k) {
      iter_seen[k] = c->h_mx[k].iter;
      resid_seen[k] = c->h_mx[k].residuals;
      all_done = all_done && c->h_mx[k].phase == 3;
    }
  }
  double spmv_ms = 0.0, resid_ms = 0.0, store_ms = 0.0;
  int spmv_ran = 0, resid_ran = 0;
  int store_pair = -1;  // the event pair of the launch that wrote the float blocks
  std::vector<int> chunk_spmv;
  while (!all_done) {
    chunk_spmv.clear();
    for (int i = 0; i < chunk; ++i) {
      if (timing) {
        if (int e = pair(spmv_ev)) return e;
        chunk_spmv.push_back(spmv_ev.back());
        HIPCHK(c, hipEventRecord(c->ev[spmv_ev.back()], c->stream));
      }
      if (c->blocks32_deferred) {  // the first inner product writes the float blocks
        pd::launch_mx_spmv_store(a, c->val32, c->stream);
        c->blocks32_deferred = false;
        if (timing) store_pair = spmv_ev.back();
      } else {
        pd::launch_mx_spmv(a, c->stream);
      }
      if (timing) HIPCHK(c, hipEventRecord(c->ev[spmv_ev.back() + 1], c->stream));
      pd::launch_mx_update(a, c->stream);
      pd::launch_mx_dir(a, c->stream);
    }
    pd::launch_mx_apply(a, c->stream);
    int rpair = -1;
    if (timing) {
      if (int e = pair(resid_ev)) return e;
      rpair = resid_ev.back();
      HIPCHK(c, hipEventRecord(c->ev[rpair], c->stream));
    }
    pd::launch_mx_resid(a, c->stream);
    if (timing) HIPCHK(c, hipEventRecord(c->ev[rpair + 1], c->stream));
    CHECK_LAUNCH(c, "mixed iteration");
    HIPCHK(c, hipMemcpyAsync(c->h_mx, c->mx, sizeof(pd_mx_state) * ns, hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    ++syncs;
    // a launch did work if any specimen needed it: the inner SpMVs that ran
    // are the first max_k (iterations of k in this chunk) of the chunk
    int ran = 0;
    bool resid_ran_now = false;
    all_done = true;
    for (int k = 0; k < ns; ++k) {
      const pd_mx_state& h = c->h_mx[k];
      ran = std::max(ran, h.iter - iter_seen[k]);
      resid_ran_now = resid_ran_now || h.residuals > resid_seen[k];
      iter_seen[k] = h.iter;
      resid_seen[k] = h.residuals;
      all_done = all_done && h.phase == 3;
    }
    if (timing) {
      ran = std::min((int)chunk_spmv.size(), ran);
      for (int k = 0; k < ran; ++k) {
        if (chunk_spmv[k] == store_pair) continue;  // reported on its own
        spmv_ms += elapsed(c->ev[chunk_spmv[k]], c->ev[chunk_spmv[k] + 1]);
        ++spmv_ran;
      }
      if (store_pair >= 0 && syncs == 1)
        store_ms = elapsed(c->ev[store_pair], c->ev[store_pair + 1]);
      if (syncs == 1 && !resumed && !fused) {
        resid_ms += elapsed(c->ev[resid_ev[0]], c->ev[resid_ev[0] + 1]);
        ++resid_ran;
      }
      if (resid_ran_now) {
        resid_ms += elapsed(c->ev[rpair], c->ev[rpair + 1]);
        ++resid_ran;
      }
    }
    if (all_done) break;
    // every chunk advances some specimen's inner iterations or refinements
    if (syncs > ns * (prm->max_iter + lim.max_refine + 2))
      return fail(c, PD_ERR_HIP, "mixed solve state did not terminate on the device");
    if (singles > 0) {
      --singles;
      chunk = 1;
    } else {
      next = std::min(next * 2, 32);
      chunk = next;
    }
  }
  int it = 0, refs = 0, worst = 0;
  bool conv = true;
  for (int k = 0; k < ns; ++k) {
    const pd_mx_state& h = c->h_mx[k];
    it = std::max(it, h.iter);
    refs = std::max(refs, h.refinements);
    conv = conv && h.converged;
    const pd_mx_state& w = c->h_mx[worst];
    if (h.rnorm * w.bnorm > w.rnorm * h.bnorm) worst = k;
  }
  c->last_mx_iters = it;
  c->last_converged = conv;
  c->last_solve_prec = 1;
  if (st) {
    st->cg_iterations = it;
    st->converged = conv ? 1 : 0;
    st->residual_norm = c->h_mx[worst].rnorm;
    st->b_norm = c->h_mx[worst].bnorm;
    st->refinements = refs;
    st->host_syncs += syncs;
    if (timing) {
      st->ms_spmv_total = spmv_ms;
      st->spmv_launches = spmv_ran;
      st->ms_resid_total = resid_ms;
      st->resid_launches = resid_ran;
      st->ms_spmv_store = store_ms;
    }
  }
  return PD_OK;
}

int enqueue_update(pd_ctx* c, const pd_step_params* prm,
                   const pd_step_consts* consts = nullptr) {
  if (int e = ensure_blocks(c)) return e;  // from the positions they belong to
  double lo[3], hi[3];
  for (int k = 0; k < 3; ++k) {
    lo[k] = prm->bmin[k] + prm->wall_radius;
    hi[k] = prm->bmax[k] - prm->wall_radius;
  }
  pd::launch_update(c->n, c->x, c->v, c->sol, c->act, prm->integrator, prm->dt, prm->walls, lo, hi,
                    prm->restitution, consts, c->stream);
  CHECK_LAUNCH(c, "update");
  c->asm_fresh = false;
  return PD_OK;
}

// The learned blocks M_p of the assembled system into pc_minv (and pc_minv32
// for the mixed solve); the solve with precond 2 applies them in place of the
// assembly's block-Jacobi inverse (minv / minv32), which stays intact.
// The learned blocks M_p of the assembled system (features, then the
// network).  fused_start: after a fused assembly the same pass also forms
// the solve's start with them -- z = M r of the residual the assembly
// formed, p = z, and the <r, z> tile partials in place of the assembly's
// block-Jacobi ones.
int apply_learned(pd_ctx* c, const pd_step_params* prm, bool fused_start = false) {
  const size_t npad = (size_t)c->nslices * 64;
  if (!c->pc_feat)
    if (int e = dalloc(c, &c->pc_feat, (size_t)std::max(c->n, 1) * PC_NIN)) return e;
  if (!c->pc_minv)
    if (int e = dalloc(c, &c->pc_minv, npad * 6)) return e;
  if (prm->precision == 1 && !c->pc_minv32)
    if (int e = dalloc(c, &c->pc_minv32, npad * 6)) return e;
  pd::PcFeatArgs fa;
  fa.n = c->n;
  fa.slice_off = c->slice_off;
  fa.top = c->top;
  fa.cap = c->cap;
  fa.col = c->col;
  fa.x = c->x;
  fa.diag = c->diag;
  fa.bond_stats = c->bond_stats;
  fa.act = c->act;
  fa.body = c->body;
  fa.mat = c->dmat;
  fa.feat = c->pc_feat;
  pd::launch_pc_features(fa, c->stream);
  pd::PcNet net;
  net.w = c->pc_w;
  net.p = c->pc_p;
  float* minv32 = prm->precision == 1 ? c->pc_minv32 : nullptr;
  if (fused_start) {
    pd::PcInferStartArgs a;
    a.n = c->n;
    a.mixed = prm->precision == 1 ? 1 : 0;
    a.feat = c->pc_feat;
    a.net = net;
    a.diag = c->diag;
    a.act = c->act;
    a.minv = c->pc_minv;
    a.minv32 = minv32;
    a.r64 = c->r;
    a.p64 = c->mf ? nullptr : c->p;  // matrix-free: p lives in the record alone
    a.rec = (c->mf || c->blocks_deferred) ? c->rec : nullptr;
    a.r32 = c->r32;
    a.p32 = c->p32;
    a.part = c->part;
    pd::launch_pc_infer_start(a, c->stream);
  } else {
    pd::launch_pc_infer(c->n, c->pc_feat, net, c->diag, c->act, c->pc_minv, minv32, c->stream);
  }
  CHECK_LAUNCH(c, "learned preconditioner");
  return PD_OK;
}

// The solve with the preconditioner choice: block Jacobi records its
// iteration count; the learned preconditioner gets at most
// learned_fallback x that many iterations (when known), after which the
// block-Jacobi inverse is restored and the solve goes on from its iterate.
// may_probe (pd_step): a precond-2 solve without a current block-Jacobi
// count (none yet, learned_probe steps since the last, or a fallback in the
// previous step) runs block Jacobi instead (the assembly formed its inverse)
// and refreshes the count.  fused: the assembly formed the block-Jacobi start.
int run_solve_pc(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st,
                 bool fused = false, bool may_probe = false) {
  c->last_learned = false;
  if (prm->precond != 2) {
    if (int e = run_solve(c, prm, st, 0, fused)) return e;
    if (prm->precond == 1) {
      c->bj_iters = prm->precision == 1 ? c->last_mx_iters : c->last_iters;
      c->pc_since_probe = 0;
    }
    return PD_OK;
  }
  if (c->mat_prec != prm->precision)
    return fail(c, PD_ERR_STATE, "the matrix was not assembled with this precision");
  if (may_probe) {
    const int period = prm->learned_probe == 0 ? 16 : prm->learned_probe;
    if (c->bj_iters == 0 || c->pc_fell_back || (period > 0 && c->pc_since_probe >= period)) {
      pd_step_params p1 = *prm;
      p1.precond = 1;
      if (int e = run_solve(c, &p1, st, 0, fused)) return e;
      c->bj_iters = prm->precision == 1 ? c->last_mx_iters : c->last_iters;
      c->pc_since_probe = 0;
      c->pc_fell_back = false;
      if (st) st->precond_probe = 1;
      return PD_OK;
    }
    ++c->pc_since_probe;
  }
  if (int e = apply_learned(c, prm, fused)) return e;
  c->last_learned = true;
  pd_step_params p2 = *prm;
  const double f = prm->learned_fallback > 0.0 ? prm->learned_fallback : 1.5;
  if (c->bj_iters > 0)
    p2.max_iter = std::min(prm->max_iter, (int32_t)ceil(f * (double)c->bj_iters));
  if (int e = run_solve(c, &p2, st, 0, fused)) return e;
  if (c->last_converged || p2.max_iter >= prm->max_iter) return PD_OK;
  // fall back to block Jacobi (the assembly's inverse, still in minv /
  // minv32) from the current iterate; the next step probes (its count is
  // then the baseline)
  c->pc_fell_back = true;
  const int used = prm->precision == 1 ? c->last_mx_iters : c->last_iters;
  pd_step_params p3 = *prm;
  p3.precond = 1;
  p3.max_iter = std::max(0, prm->max_iter - used);
  p3.inner_max_iter = prm->inner_max_iter > 0 ? prm->inner_max_iter : prm->max_iter;
  pd_step_stats s3 = {};
  if (int e = run_solve(c, &p3, st ? &s3 : nullptr)) return e;
  if (st) {
    st->cg_iterations = used + s3.cg_iterations;
    st->converged = s3.converged;
    st->residual_norm = s3.residual_norm;
    st->b_norm = s3.b_norm;
    st->refinements += s3.refinements;
    st->host_syncs += s3.host_syncs;
    st->ms_spmv_total += s3.ms_spmv_total;
    st->spmv_launches += s3.spmv_launches;
    st->ms_resid_total += s3.ms_resid_total;
    st->resid_launches += s3.resid_launches;
    st->precond_fallbacks = 1;
  }
  return PD_OK;
}

// ---------------------------------------------------------------------------
// Captured steps (graph = 1).  One hipGraph holds the assembly, the stretch
// commit (a copy, so the captured pointers keep their roles), the solve's
// start and K iterations (mixed: K inner iterations, the correction and the
// residual test), and the pinned copies of the solver states and counters;
// a second graph holds the position update.  Step scalars (dt, clock) are
// read from device memory, so the graphs are replayed every step.  If a
// specimen is still iterating after the graph, the host-driven loop goes on
// from the device state before the update graph runs.
// ---------------------------------------------------------------------------
void destroy_graphs(pd_ctx* c) {
  if (c->g_solve) (void)hipGraphExecDestroy(c->g_solve);
  if (c->g_update) (void)hipGraphExecDestroy(c->g_update);
  c->g_solve = c->g_update = nullptr;
  c->graph_key.clear();
}

std::string graph_key(pd_ctx* c, const pd_step_params* prm, int K) {
  const MixedLimits lim = mixed_limits(prm);
  char buf[1024];
  snprintf(buf, sizeof(buf),
           "%d %d %d %d %a %d %a %d %d %d %a %a %a %a %a %a %a %a %d %d %d %p %p %p %p %p %p %p %p",
           prm->matrix_free, prm->precision, prm->precond, prm->integrator, prm->rtol,
           prm->max_iter,
           lim.inner_rtol, lim.inner_max, lim.max_refine, prm->walls, prm->bmin[0],
           prm->bmin[1], prm->bmin[2], prm->bmax[0], prm->bmax[1], prm->bmax[2],
           prm->wall_radius, prm->restitution, c->unroll[prm->precision], c->nspec, K,
           (void*)c->s, (void*)c->s_new, (void*)c->val, (void*)c->col, (void*)c->val32,
           (void*)c->tile_spec, (void*)c->st, (void*)c->rec);
  return std::string(buf);
}

int capture(pd_ctx* c, hipGraphExec_t* out, const std::function<int()>& body) {
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  const int e = body();
  const hipError_t ee = hipStreamEndCapture(c->stream, &g);
  if (e != PD_OK) {
    if (g) (void)hipGraphDestroy(g);
    return e;
  }
  if (ee != hipSuccess || !g)
    return fail(c, PD_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(ee));
  const hipError_t ei = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess)
    return fail(c, PD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  return PD_OK;
}

int capture_step(pd_ctx* c, const pd_step_params* prm, int K) {
  destroy_graphs(c);
  const int ns = c->nspec;
  if (!c->d0_valid) {  // once, outside the captured step
    pd::launch_bond_d0(c->n, c->slice_off, c->cap, c->col, c->x0, c->d0, c->d0_bad, c->stream);
    CHECK_LAUNCH(c, "reference bond lengths");
    c->d0_valid = true;
  }
  // no allocation inside the capture
  if (prm->matrix_free || defer_blocks(prm) || prm->precision == 1)
    if (int e = ensure_rec(c)) return e;
  if (prm->precision == 1)
    if (int e = ensure_mixed(c)) return e;
  int e = capture(c, &c->g_solve, [&]() -> int {
    const bool fuse = true;  // precond 0 or 1 (check_params)
    if (int e2 = enqueue_assemble(c, prm, c->consts, fuse)) return e2;
    pd::launch_copy(c->s, c->s_new, c->n, c->stream);  // core/simulator.cuh:2612
    if (prm->precision == 1) {
      const pd::MxArgs a = mx_args(c, prm->precond);
      const MixedLimits lim = mixed_limits(prm);
      pd::launch_mx_setup(c->mx, ns, prm->rtol, prm->max_iter, lim.inner_rtol, lim.inner_max,
                          lim.max_refine, c->stream);
      if (fuse) {
        pd::launch_mx_resid_fold(a, c->stream);
      } else {
        pd::launch_mx_rec_x(a, c->stream);
        pd::launch_mx_resid(a, c->stream);
      }
      for (int i = 0; i < K; ++i) {
        if (c->blocks32_deferred) {  // the first inner product writes the float blocks
          pd::launch_mx_spmv_store(a, c->val32, c->stream);
          c->blocks32_deferred = false;
        } else {
          pd::launch_mx_spmv(a, c->stream);
        }
        pd::launch_mx_update(a, c->stream);
        pd::launch_mx_dir(a, c->stream);
      }
      pd::launch_mx_apply(a, c->stream);
      pd::launch_mx_resid(a, c->stream);
      HIPCHK(c, hipMemcpyAsync(c->h_mx, c->mx, sizeof(pd_mx_state) * ns, hipMemcpyDeviceToHost,
                               c->stream));
    } else {
      const pd::CgArgs a = cg_args(c, prm->precond);
      if (fuse) {  // the fold sets the solver state up too
        pd::launch_cg_init_fold(a, prm->rtol, prm->max_iter, c->stream);
      } else if (c->blocks_deferred) {
        pd::launch_cg_setup(c->st, ns, prm->rtol, prm->max_iter, c->stream);
        pd::launch_cg_init_store(a, c->stream);
        c->blocks_deferred = false;
      } else {
        pd::launch_cg_setup(c->st, ns, prm->rtol, prm->max_iter, c->stream);
        pd::launch_cg_init(a, c->stream);
      }
      for (int i = 0; i < K; ++i) {
        if (c->blocks_deferred) {  // the first product writes the blocks
          pd::launch_cg_spmv_store(a, c->stream);
          c->blocks_deferred = false;
        } else {
          pd::launch_cg_spmv(a, c->stream);
        }
        pd::launch_cg_update(a, i, c->stream);
        pd::launch_cg_dir(a, i, c->stream);
      }
      HIPCHK(c, hipMemcpyAsync(c->h_st, c->st, sizeof(pd_cg_state) * ns, hipMemcpyDeviceToHost,
                               c->stream));
    }
    return PD_OK;
  });
  if (e == PD_OK) e = capture(c, &c->g_update, [&]() { return enqueue_update(c, prm, c->consts); });
  if (e != PD_OK) {
    destroy_graphs(c);
    return e;
  }
  c->graph_key = graph_key(c, prm, K);
  c->graph_k = K;
  return PD_OK;
}

int run_step_graph(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st) {
  if (prm->contact_radius > 0.0)
    return fail(c, PD_ERR_ARG, "captured steps (graph = 1) need contact off");
  if (prm->track_gap) return fail(c, PD_ERR_ARG, "captured steps (graph = 1) do not track gaps");
  if (prm->precision == 1)
    if (int e = ensure_mixed(c)) return e;
  if (!c->consts) {
    if (int e = dalloc(c, &c->consts, 1)) return e;
    if (hipHostMalloc((void**)&c->h_consts, sizeof(pd_step_consts), hipHostMallocDefault) !=
        hipSuccess)
      return fail(c, PD_ERR_HIP, "hipHostMalloc failed");
  }
  const int last = prm->precision == 1 ? c->last_mx_iters : c->last_iters;
  int K = c->graph_k;
  if (!c->g_solve || last + 1 > K || last + 4 < K) K = std::max(2, last + 1);
  if (graph_key(c, prm, K) != c->graph_key)
    if (int e = capture_step(c, prm, K)) return e;
  const bool timing = prm->timing != 0;
  if (timing)
    if (int e = ensure_events(c, 8)) return e;
  c->h_consts->dt = prm->dt;
  c->h_consts->dt2 = prm->integrator == 1 ? (prm->dt * prm->dt) / 4.0 : prm->dt * prm->dt;
  c->h_consts->clock = prm->clock_scale;
  c->h_consts->gravity = prm->gravity;
  HIPCHK(c, hipMemcpyAsync(c->consts, c->h_consts, sizeof(pd_step_consts), hipMemcpyHostToDevice,
                           c->stream));
  if (timing) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  HIPCHK(c, hipGraphLaunch(c->g_solve, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mat_prec = prm->precision;
  c->asm_fresh = true;  // the graph's assembly ran
  if (int e = run_solve(c, prm, st, K)) return e;
  read_assemble_stats(c, prm, st);
  if (timing) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipGraphLaunch(c->g_update, c->stream));
  c->asm_fresh = false;
  if (timing) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (st) {
    st->host_syncs += 1;
    if (timing) {
      st->ms_solve = elapsed(c->ev[0], c->ev[1]);  // assembly + solve
      st->ms_update = elapsed(c->ev[1], c->ev[2]);
    }
  }
  if (!c->last_converged)
    return fail(c, PD_ERR_NOT_CONVERGED, "CG reached the iteration limit before rtol");
  return PD_OK;
}

// matrix_free = 2: the operator of this step, 1 (recompute the blocks in
// every product) or 0 (the first product stores them, the others stream
// them).  Both give identical results, so the choice is by cost alone: with
// n the previous solve's iteration count, recompute when n t_mf <= t_store
// + (n - 1) t_st.  Until both have been measured on this context the steps
// measure them in turn (recomputing first); captured steps, which cannot
// time their products, recompute until then.  Mixed precision streams its
// float blocks (0).
int auto_operator(pd_ctx* c, const pd_step_params* prm) {
  if (prm->precision != 0) return 0;
  if (c->op_t_mf <= 0.0) return 1;
  if (c->op_t_store <= 0.0) return prm->graph ? 1 : 0;
  const double n = (double)std::max(1, c->last_iters);
  return n * c->op_t_mf <= c->op_t_store + (n - 1.0) * c->op_t_st ? 1 : 0;
}

void reset_operator_times(pd_ctx* c) { c->op_t_mf = c->op_t_store = c->op_t_st = 0.0; }

void zero_timing(pd_step_stats* st) {
  if (!st) return;
  st->ms_contact = st->ms_assemble = st->ms_solve = st->ms_update = 0.0;
  st->ms_spmv_total = 0.0;
  st->ms_spmv_store = 0.0;
  st->spmv_launches = 0;
  st->host_syncs = 0;
  st->refinements = 0;
  st->ms_resid_total = 0.0;
  st->resid_launches = 0;
  st->precond_fallbacks = 0;
  st->precond_probe = 0;
  st->operator_mf = 0;
  st->contact_grid = 0;
}

// One step (contact, assembly, stretch commit, solve, update) with the
// operator prm->matrix_free (0 or 1) names.
int run_step(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st) {
  if (prm->graph) {
    const int e = run_step_graph(c, prm, st);
    if (st) st->operator_mf = c->mf;
    return e;
  }
  const bool timing = prm->timing != 0;
  if (timing)
    if (int e = ensure_events(c, 8)) return e;
  if (timing) HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  if (prm->contact_radius > 0.0) {
    if (int e = enqueue_contact(c, prm)) return e;
    if (st) st->contact_grid = c->last_grid;
  }
  if (timing) HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  if (int e = enqueue_assemble(c, prm, nullptr, true)) return e;
  if (timing) HIPCHK(c, hipEventRecord(c->ev[2], c->stream));
  std::swap(c->s, c->s_new);  // core/simulator.cuh:2612
  if (int e = run_solve_pc(c, prm, st, true, true)) return e;
  read_assemble_stats(c, prm, st);  // the solve's waits covered the copies
  if (timing) HIPCHK(c, hipEventRecord(c->ev[3], c->stream));
  if (int e = enqueue_update(c, prm)) return e;
  if (timing) HIPCHK(c, hipEventRecord(c->ev[4], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (st && timing) {
    st->ms_contact = elapsed(c->ev[0], c->ev[1]);
    st->ms_assemble = elapsed(c->ev[1], c->ev[2]);
    st->ms_solve = elapsed(c->ev[2], c->ev[3]);
    st->ms_update = elapsed(c->ev[3], c->ev[4]);
  }
  if (st) st->operator_mf = c->mf;
  if (!c->last_converged)
    return fail(c, PD_ERR_NOT_CONVERGED, "CG reached the iteration limit before rtol");
  return PD_OK;
}


}  // namespace

extern "C" {

const char* pd_last_error(const pd_ctx* ctx) {
  if (ctx) return ctx->err.c_str();
  return g_last_error.c_str();
}

int pd_ctx_create(const pd_config* cfg, pd_ctx** out) {
  if (!out) return fail(nullptr, PD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (!cfg) return fail(nullptr, PD_ERR_ARG, "null config");
  if (cfg->num_particles < 0) return fail(nullptr, PD_ERR_ARG, "num_particles < 0");
  if (cfg->max_bonds < 1 || cfg->max_bonds > PD_MAX_BONDS_LIMIT)
    return fail(nullptr, PD_ERR_ARG, "max_bonds outside [1, 1024]");
  if (cfg->num_bodies < 1 || cfg->num_bodies > PD_MAX_BODIES)
    return fail(nullptr, PD_ERR_ARG, "num_bodies outside [1, 8]");
  for (int i = 0; i < cfg->num_bodies; ++i) {
    const pd_material& m = cfg->material[i];
    if (!(m.horizon > 0.0) || !(m.radius > 0.0) || !(m.mass > 0.0))
      return fail(nullptr, PD_ERR_ARG, "material horizon, radius and mass must be > 0");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, PD_ERR_NO_DEVICE, "no HIP device available");
  if (cfg->device < 0 || cfg->device >= ndev)
    return fail(nullptr, PD_ERR_ARG, "device ordinal out of range");
  pd_ctx* c = new pd_ctx();
  c->device = cfg->device;
  c->n = cfg->num_particles;
  c->nslices = (c->n + 63) / 64;
  c->max_bonds = cfg->max_bonds;
  c->nbody = cfg->num_bodies;
  for (int i = 0; i < PD_MAX_BODIES; ++i) c->mat[i] = cfg->material[i];
  auto bail = [&](int code) {
    std::string msg = c->err;
    pd_ctx_destroy(c);
    g_last_error = msg;
    return code;
  };
  if (hipSetDevice(c->device) != hipSuccess) return bail(fail(c, PD_ERR_HIP, "hipSetDevice failed"));
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(c, PD_ERR_HIP, "hipStreamCreate failed"));
  const size_t n = (size_t)c->n;
  const size_t npad = (size_t)c->nslices * 64;
  const size_t ntiles = (n + PD_BLOCK - 1) / PD_BLOCK;
  int e = PD_OK;
#define ALLOC(ptr, count) \
  if ((e = dalloc(c, &(ptr), (count))) != PD_OK) return bail(e)
  ALLOC(c->x, n);
  ALLOC(c->v, n);
  ALLOC(c->x0, n);
  ALLOC(c->sol, n);
  ALLOC(c->f, n);
  ALLOC(c->b, n);
  ALLOC(c->fc, n);
  ALLOC(c->vbc, n);
  ALLOC(c->s, n);
  ALLOC(c->s_new, n);
  ALLOC(c->s00, n);
  ALLOC(c->act, n);
  ALLOC(c->body, n);
  ALLOC(c->bc, n);
  ALLOC(c->top, n);
  ALLOC(c->cap, n);
  ALLOC(c->slice_off, (size_t)c->nslices + 1);
  ALLOC(c->diag, npad * 6);
  ALLOC(c->minv, npad * 6);
  ALLOC(c->r, n);
  ALLOC(c->z, n);
  ALLOC(c->p, n);
  ALLOC(c->y, n);
  ALLOC(c->part, 4 * ntiles);
  ALLOC(c->st, 1);
  c->state_cap = 1;
  ALLOC(c->counters, 1);
  ALLOC(c->d0_bad, 1);
  ALLOC(c->arec, n * 8);
  ALLOC(c->dmat, PD_MAX_BODIES);
  ALLOC(c->dhorizon, PD_MAX_BODIES);
  ALLOC(c->gap_val, n);
  ALLOC(c->gap_q, n);
  ALLOC(c->gpart_v, ntiles);
  ALLOC(c->gpart_p, ntiles);
  ALLOC(c->gpart_q, ntiles);
  ALLOC(c->gap_out, 1);
  ALLOC(c->gap_pq, 2);
  ALLOC(c->keys_a, n);
  ALLOC(c->keys_b, n);
  ALLOC(c->idx_a, n);
  ALLOC(c->idx_b, n);
  ALLOC(c->hist, pd::radix_hist_words(c->n));
  ALLOC(c->bbox_part, 1024 * 6 * PD_MAX_BODIES);
  ALLOC(c->bbox, 6 * PD_MAX_BODIES);
  ALLOC(c->plan, 1);
  ALLOC(c->counts, n);
  ALLOC(c->zone_cnt, pd::zone_tile_words(c->n));
  ALLOC(c->cand, n);
  ALLOC(c->widths, (size_t)c->nslices);
  ALLOC(c->bond_stats, n);
  ALLOC(c->x_bak, n);
  ALLOC(c->v_bak, n);
  ALLOC(c->top_bak, n);
#undef ALLOC
  if (hipHostMalloc((void**)&c->h_st, sizeof(pd_cg_state), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_cnt, sizeof(pd_step_counters), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_bbox, 6 * PD_MAX_BODIES * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_gap, sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_gap_pq, 2 * sizeof(int32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_zone_count, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_plan, sizeof(pd::ContactPlan), hipHostMallocDefault) != hipSuccess)
    return bail(fail(c, PD_ERR_HIP, "hipHostMalloc failed"));
  pd_mat_dev hm[PD_MAX_BODIES];
  double hh[PD_MAX_BODIES];
  for (int i = 0; i < PD_MAX_BODIES; ++i) {
    hm[i].horizon = c->mat[i].horizon;
    hm[i].radius = c->mat[i].radius;
    hm[i].c_bond = c->mat[i].c_bond;
    hm[i].mass = c->mat[i].mass;
    hh[i] = c->mat[i].horizon;
  }
  if (hipMemcpyAsync(c->dmat, hm, sizeof(hm), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->dhorizon, hh, sizeof(hh), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return bail(fail(c, PD_ERR_HIP, "material upload failed"));
  *out = c;
  return PD_OK;
}

int pd_ctx_destroy(pd_ctx* c) {
  if (!c) return PD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->x); dfree(c->v); dfree(c->x0); dfree(c->sol); dfree(c->f); dfree(c->b);
  dfree(c->fc); dfree(c->vbc); dfree(c->s); dfree(c->s_new); dfree(c->s00);
  dfree(c->act); dfree(c->body); dfree(c->bc); dfree(c->top); dfree(c->cap);
  dfree(c->slice_off); dfree(c->col); dfree(c->val); dfree(c->d0); dfree(c->diag); dfree(c->minv);
  dfree(c->r); dfree(c->z); dfree(c->p); dfree(c->y); dfree(c->part); dfree(c->st);
  dfree(c->counters); dfree(c->d0_bad); dfree(c->arec); dfree(c->dmat); dfree(c->dhorizon); dfree(c->rec);
  dfree(c->gap_val); dfree(c->gap_q); dfree(c->gpart_v); dfree(c->gpart_p); dfree(c->gpart_q);
  dfree(c->gap_out); dfree(c->gap_pq);
  dfree(c->keys_a); dfree(c->keys_b); dfree(c->idx_a); dfree(c->idx_b); dfree(c->hist);
  dfree(c->bbox_part); dfree(c->bbox); dfree(c->plan); dfree(c->cell_start); dfree(c->cell_end);
  dfree(c->counts); dfree(c->widths); dfree(c->zone_cnt); dfree(c->cand); dfree(c->x_bak); dfree(c->v_bak); dfree(c->top_bak);
  dfree(c->val32); dfree(c->diag32); dfree(c->minv32); dfree(c->r32); dfree(c->z32);
  dfree(c->p32); dfree(c->y32); dfree(c->d32); dfree(c->mx);
  dfree(c->tile_spec); dfree(c->spec_tile0); dfree(c->spec_ntiles);
  destroy_graphs(c);
  dfree(c->consts);
  if (c->h_consts) (void)hipHostFree(c->h_consts);
  dfree(c->bond_stats); dfree(c->pc_feat); dfree(c->pc_p); dfree(c->pc_w);
  dfree(c->pc_minv); dfree(c->pc_minv32);
  if (c->h_mx) (void)hipHostFree(c->h_mx);
  if (c->h_st) (void)hipHostFree(c->h_st);
  if (c->h_cnt) (void)hipHostFree(c->h_cnt);
  if (c->h_bbox) (void)hipHostFree(c->h_bbox);
  if (c->h_gap) (void)hipHostFree(c->h_gap);
  if (c->h_gap_pq) (void)hipHostFree(c->h_gap_pq);
  if (c->h_zone_count) (void)hipHostFree(c->h_zone_count);
  if (c->h_plan) (void)hipHostFree(c->h_plan);
  for (hipEvent_t e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PD_OK;
}

int pd_upload(pd_ctx* c, int32_t field, const void* src, size_t bytes) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  FieldInfo fi = field_info(c, field);
  if (!fi.dptr || !fi.writable) return fail(c, PD_ERR_ARG, "field cannot be uploaded");
  if (bytes != fi.elem * (size_t)c->n) return fail(c, PD_ERR_ARG, "upload size mismatch");
  if (c->n == 0) return PD_OK;
  if (!src) return fail(c, PD_ERR_ARG, "null source");
  if (field == PD_FIELD_POSITION || field == PD_FIELD_BOND_TOP || field == PD_FIELD_BODY) {
    // blocks a solve has not written yet belong to the state being replaced
    HIPCHK(c, hipSetDevice(c->device));
    if (int e = ensure_blocks(c)) return e;
  }
  if (field == PD_FIELD_BOND_TOP) {
    if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built");
    const int32_t* t = (const int32_t*)src;
    for (int32_t i = 0; i < c->n; ++i)
      if (t[i] < -1 || t[i] >= c->h_cap[i])
        return fail(c, PD_ERR_ARG, "bond top outside [-1, capacity)");
  }
  if (field == PD_FIELD_ORIGINAL_POSITION) c->d0_valid = false;
  c->asm_fresh = false;
  if (field == PD_FIELD_BODY) {
    const int32_t* bt = (const int32_t*)src;
    for (int32_t i = 0; i < c->n; ++i)
      if (bt[i] < 0 || bt[i] >= c->nbody)
        return fail(c, PD_ERR_ARG, "body index outside [0, num_bodies)");
  }
  HIPCHK(c, hipSetDevice(c->device));
  if (fi.vec4) {
    // the w lane is not part of the state: store it as 0
    std::vector<double> tmp((const double*)src, (const double*)src + 4 * (size_t)c->n);
    for (size_t i = 0; i < (size_t)c->n; ++i) tmp[4 * i + 3] = 0.0;
    HIPCHK(c, hipMemcpyAsync(*fi.dptr, tmp.data(), bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    HIPCHK(c, hipMemcpyAsync(*fi.dptr, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return PD_OK;
}

int pd_download(pd_ctx* c, int32_t field, void* dst, size_t bytes) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  FieldInfo fi = field_info(c, field);
  if (!fi.dptr) return fail(c, PD_ERR_ARG, "unknown field");
  if (bytes != fi.elem * (size_t)c->n) return fail(c, PD_ERR_ARG, "download size mismatch");
  if (c->n == 0) return PD_OK;
  if (!dst) return fail(c, PD_ERR_ARG, "null destination");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipMemcpyAsync(dst, *fi.dptr, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_build_bonds(pd_ctx* c, pd_bond_stats* stats) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(c, hipEventCreate(&e0));
  HIPCHK(c, hipEventCreate(&e1));
  HIPCHK(c, hipEventRecord(e0, c->stream));
  const int32_t n = c->n;
  std::vector<int32_t> widths(c->nslices, 0);
  std::vector<int32_t> counts(n, 0);
  if (n > 0) {
    double hmax = 0.0;
    for (int i = 0; i < c->nbody; ++i) hmax = std::max(hmax, c->mat[i].horizon);
    // a cell slightly wider than the horizon keeps every bonded pair within
    // one cell of each other despite rounding in the cell index
    pd::GridView g;
    if (int e = build_grid(c, hmax * (1.0 + 1e-7), &g)) return e;
    pd::launch_bond_count(n, c->x, c->body, c->act, c->dhorizon, g, c->counts, c->stream);
    CHECK_LAUNCH(c, "bond count");
    pd::launch_slice_width(c->counts, n, c->widths, c->stream);
    CHECK_LAUNCH(c, "slice width");
    HIPCHK(c, hipMemcpyAsync(widths.data(), c->widths, sizeof(int32_t) * c->nslices,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(counts.data(), c->counts, sizeof(int32_t) * n,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int32_t mx = *std::max_element(widths.begin(), widths.end());
    if (mx > c->max_bonds) {
      c->bonds_ready = false;
      return fail(c, PD_ERR_BOND_OVERFLOW,
                  "a particle has " + std::to_string(mx) + " bonds, more than max_bonds = " +
                      std::to_string(c->max_bonds));
    }
    if (int e = set_slice_offsets(c, widths)) return e;
    pd::launch_bond_fill(n, c->x, c->body, c->act, c->dhorizon, g, c->counts, c->slice_off, c->col,
                         c->top, c->cap, c->stream);
    CHECK_LAUNCH(c, "bond fill");
  } else {
    if (int e = set_slice_offsets(c, widths)) return e;
  }
  HIPCHK(c, hipEventRecord(e1, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int64_t total = 0;
  int32_t mn = n ? counts[0] : 0, mxc = 0;
  for (int32_t k : counts) {
    total += k;
    mn = std::min(mn, k);
    mxc = std::max(mxc, k);
  }
  c->live_bonds = total;
  c->bonds_ready = true;
  c->mat_prec = -1;
  c->last_iters = 0;
  c->h_cap = counts;
  c->d0_valid = false;
  c->blocks_deferred = false;  // new bond lists: no matrix to keep
  c->blocks32_deferred = false;
  c->asm_fresh = false;
  reset_operator_times(c);
  if (stats) {
    stats->total_bonds = total;
    stats->min_bonds = mn;
    stats->max_bonds = mxc;
    stats->slots = c->slots * 64;
    stats->ms_build = elapsed(e0, e1);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return PD_OK;
}

int pd_set_bonds(pd_ctx* c, const int32_t* counts, const int32_t* lists, int32_t stride) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  const int32_t n = c->n;
  if (n > 0 && (!counts || !lists)) return fail(c, PD_ERR_ARG, "null bond arrays");
  if (stride < 0) return fail(c, PD_ERR_ARG, "negative stride");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int32_t> widths(c->nslices, 0);
  for (int32_t p = 0; p < n; ++p) {
    if (counts[p] < 0 || counts[p] > stride) return fail(c, PD_ERR_ARG, "count outside [0, stride]");
    if (counts[p] > c->max_bonds) return fail(c, PD_ERR_BOND_OVERFLOW, "row exceeds max_bonds");
    for (int32_t k = 0; k < counts[p]; ++k) {
      const int32_t q = lists[(int64_t)p * stride + k];
      if (q < 0 || q >= n || q == p) return fail(c, PD_ERR_ARG, "bond index out of range");
    }
    widths[p >> 6] = std::max(widths[p >> 6], counts[p]);
  }
  if (int e = set_slice_offsets(c, widths)) return e;
  std::vector<int32_t> col((size_t)c->slots * 64, 0), top(n), cap(n);
  int64_t total = 0;
  for (int32_t p = 0; p < n; ++p) {
    const int64_t base = c->h_slice_off[p >> 6];
    for (int32_t k = 0; k < counts[p]; ++k)
      col[(size_t)(base + k) * 64 + (p & 63)] = lists[(int64_t)p * stride + k];
    top[p] = counts[p] - 1;
    cap[p] = counts[p];
    total += counts[p];
  }
  if (n > 0) {
    HIPCHK(c, hipMemcpyAsync(c->col, col.data(), col.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->top, top.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(c->cap, cap.data(), n * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->live_bonds = total;
  c->bonds_ready = true;
  c->mat_prec = -1;
  c->last_iters = 0;
  c->h_cap = cap;
  c->d0_valid = false;
  c->blocks_deferred = false;  // new bond lists: no matrix to keep
  c->blocks32_deferred = false;
  c->asm_fresh = false;
  reset_operator_times(c);
  return PD_OK;
}

int pd_get_bonds(pd_ctx* c, int32_t* capacity, int32_t* lists, int32_t stride) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built");
  const int32_t n = c->n;
  if (n == 0) return PD_OK;
  if (!capacity || !lists) return fail(c, PD_ERR_ARG, "null output");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<int32_t> col((size_t)c->slots * 64);
  HIPCHK(c, hipMemcpyAsync(capacity, c->cap, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
  if (!col.empty())
    HIPCHK(c, hipMemcpyAsync(col.data(), c->col, col.size() * sizeof(int32_t),
                             hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int32_t p = 0; p < n; ++p) {
    if (capacity[p] > stride) return fail(c, PD_ERR_ARG, "stride smaller than a row");
    const int64_t base = c->h_slice_off[p >> 6];
    for (int32_t k = 0; k < capacity[p]; ++k)
      lists[(int64_t)p * stride + k] = col[(size_t)(base + k) * 64 + (p & 63)];
  }
  return PD_OK;
}

int pd_get_matrix(pd_ctx* c, double* diag9, double* vals9, int32_t stride) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built");
  if (c->mf)
    return fail(c, PD_ERR_STATE, "the last assembly was matrix-free: bond blocks not stored");
  const int32_t n = c->n;
  if (n == 0) return PD_OK;
  if (!diag9 || !vals9) return fail(c, PD_ERR_ARG, "null output");
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = ensure_blocks(c)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> dg((size_t)c->nslices * 64 * 6), vv((size_t)c->slots * 64 * 6);
  std::vector<int32_t> top(n);
  HIPCHK(c, hipMemcpyAsync(dg.data(), c->diag, dg.size() * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  std::vector<float> vf;
  if (!vv.empty()) {
    if (c->mat_prec == 1) {
      vf.resize(vv.size());
      HIPCHK(c, hipMemcpyAsync(vf.data(), c->val32, vf.size() * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
    } else {
      HIPCHK(c, hipMemcpyAsync(vv.data(), c->val, vv.size() * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPCHK(c, hipMemcpyAsync(top.data(), c->top, n * sizeof(int32_t), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (size_t k = 0; k < vf.size(); ++k) vv[k] = (double)vf[k];
  auto unpack = [](const double* six, size_t stridev, double* out9) {
    const double a00 = six[0], a01 = six[stridev], a02 = six[2 * stridev],
                 a11 = six[3 * stridev], a12 = six[4 * stridev], a22 = six[5 * stridev];
    const double m[9] = {a00, a01, a02, a01, a11, a12, a02, a12, a22};
    memcpy(out9, m, sizeof(m));
  };
  for (int32_t p = 0; p < n; ++p) {
    const int s = p >> 6, lane = p & 63;
    unpack(dg.data() + (size_t)s * 384 + lane, 64, diag9 + 9 * (size_t)p);
    if (top[p] + 1 > stride) return fail(c, PD_ERR_ARG, "stride smaller than a row");
    const int64_t base = c->h_slice_off[s];
    for (int32_t k = 0; k <= top[p]; ++k) {
      double* out = vals9 + 9 * ((size_t)p * stride + k);
      if (c->mat_prec == 1) {  // float: a plane of 4 values per lane, then one of 2
        const double* sl = vv.data() + (size_t)(base + k) * 384;
        const double six[6] = {sl[4 * lane], sl[4 * lane + 1], sl[4 * lane + 2],
                               sl[4 * lane + 3], sl[256 + 2 * lane], sl[256 + 2 * lane + 1]};
        unpack(six, 1, out);
      } else {  // FP64 planes of pairs: value i at (i / 2) * 128 + 2 lane + i % 2
        const double* sl = vv.data() + (size_t)(base + k) * 384 + 2 * lane;
        const double six[6] = {sl[0], sl[1], sl[128], sl[129], sl[256], sl[257]};
        unpack(six, 1, out);
      }
    }
  }
  return PD_OK;
}

int pd_assemble(pd_ctx* c, const pd_step_params* prm_in, pd_step_stats* st) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (int e = check_params(c, prm_in)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  zero_timing(st);
  pd_step_params q = *prm_in;
  if (q.matrix_free == 2) q.matrix_free = auto_operator(c, &q);
  const pd_step_params* prm = &q;
  if (prm->contact_radius > 0.0) {
    if (int e = enqueue_contact(c, prm)) return e;
    if (st) st->contact_grid = c->last_grid;
  }
  if (int e = enqueue_assemble(c, prm)) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  read_assemble_stats(c, prm, st);
  return PD_OK;
}

int pd_commit_stretch(pd_ctx* c) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  std::swap(c->s, c->s_new);
  return PD_OK;
}

int pd_solve(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (int e = check_params(c, prm)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = run_solve_pc(c, prm, st)) return e;
  if (!c->last_converged)
    return fail(c, PD_ERR_NOT_CONVERGED, "CG reached the iteration limit before rtol");
  return PD_OK;
}

int pd_update(pd_ctx* c, const pd_step_params* prm) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (int e = check_params(c, prm)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = enqueue_update(c, prm)) return e;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_step(pd_ctx* c, const pd_step_params* prm, pd_step_stats* st) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (int e = check_params(c, prm)) return e;
  HIPCHK(c, hipSetDevice(c->device));
  zero_timing(st);
  if (prm->matrix_free != 2) return run_step(c, prm, st);
  // the per-step operator choice; a step that measures an operator times
  // its products (timing = 1)
  pd_step_params q = *prm;
  pd_step_stats local;
  pd_step_stats* s = st ? st : &local;
  zero_timing(s);
  q.matrix_free = auto_operator(c, prm);
  const bool measure = q.precision == 0 && !q.graph &&
                       (q.matrix_free ? c->op_t_mf <= 0.0 : c->op_t_store <= 0.0);
  if (measure) q.timing = 1;
  const int e = run_step(c, &q, s);
  if (measure && (e == PD_OK || e == PD_ERR_NOT_CONVERGED)) {
    if (q.matrix_free && s->spmv_launches > 0) {
      c->op_t_mf = s->ms_spmv_total / s->spmv_launches;
    } else if (!q.matrix_free && s->ms_spmv_store > 0.0) {
      c->op_t_store = s->ms_spmv_store;
      c->op_t_st = s->spmv_launches > 0 ? s->ms_spmv_total / s->spmv_launches : s->ms_spmv_store;
    }
  }
  return e;
}

int pd_spmv(pd_ctx* c, const double* x4, double* y4) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built");
  if (c->n == 0) return PD_OK;
  if (!x4 || !y4) return fail(c, PD_ERR_ARG, "null vector");
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = ensure_blocks(c)) return e;
  const size_t bytes = (size_t)c->n * 32;
  HIPCHK(c, hipMemcpyAsync(c->z, x4, bytes, hipMemcpyHostToDevice, c->stream));
  if (c->mat_prec == 1)
    pd::launch_spmv_wide(mx_args(c, 1), c->z, c->y, c->stream);
  else
    pd::launch_spmv_plain(spmv_view(c), c->z, c->y, c->stream);
  CHECK_LAUNCH(c, "spmv");
  HIPCHK(c, hipMemcpyAsync(y4, c->y, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_backup(pd_ctx* c) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = (size_t)c->n;
  if (n == 0) return PD_OK;
  HIPCHK(c, hipMemcpyAsync(c->top_bak, c->top, n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->x_bak, c->x, n * 32, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->v_bak, c->v, n * 32, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_restore(pd_ctx* c) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  const size_t n = (size_t)c->n;
  if (n == 0) return PD_OK;
  if (int e = ensure_blocks(c)) return e;  // from the state being replaced
  HIPCHK(c, hipMemcpyAsync(c->top, c->top_bak, n * 4, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->x, c->x_bak, n * 32, hipMemcpyDeviceToDevice, c->stream));
  c->asm_fresh = false;
  HIPCHK(c, hipMemcpyAsync(c->v, c->v_bak, n * 32, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_synchronize(pd_ctx* c) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return PD_OK;
}

int pd_spmv_bytes(pd_ctx* c, double* bytes) {
  if (!c || !bytes) return fail(c, PD_ERR_ARG, "null argument");
  if (c->mf) {  // columns, reference lengths; diagonal, top, positions, body, x, y
    *bytes = (double)c->live_bonds * 12.0 + (double)c->n * (48.0 + 4.0 + 24.0 + 4.0 + 48.0);
    return PD_OK;
  }
  const double w = c->mat_prec == 1 ? 4.0 : 8.0;  // bytes per value and vector component
  *bytes = (double)c->live_bonds * (6.0 * w + 4.0) + (double)c->n * (6.0 * w + 4.0 + 2.0 * 3.0 * w);
  return PD_OK;
}

int pd_check_division(pd_ctx* c, int64_t samples, uint64_t seed, int64_t* mismatches) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!mismatches) return fail(c, PD_ERR_ARG, "null output");
  if (samples < 0 || samples > ((int64_t)1 << 36)) return fail(c, PD_ERR_ARG, "samples outside [0, 2^36]");
  HIPCHK(c, hipSetDevice(c->device));
  unsigned long long* bad = nullptr;
  if (int e = dalloc(c, &bad, 1)) return e;
  unsigned long long h = 0;
  hipError_t err = hipMemcpyAsync(bad, &h, sizeof(h), hipMemcpyHostToDevice, c->stream);
  if (err == hipSuccess) {
    pd::launch_check_division(samples, seed, bad, c->stream);
    err = hipGetLastError();
  }
  if (err == hipSuccess) err = hipMemcpyAsync(&h, bad, sizeof(h), hipMemcpyDeviceToHost, c->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(c->stream);
  (void)hipFree(bad);
  if (err != hipSuccess) return fail(c, PD_ERR_HIP, std::string("division check: ") + hipGetErrorString(err));
  *mismatches = (int64_t)h;
  return PD_OK;
}

int pd_bench_spmv(pd_ctx* c, int32_t iters, double* ms_per_launch, double* bytes_per_launch) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!c->bonds_ready) return fail(c, PD_ERR_STATE, "bonds not built");
  if (iters < 1) return fail(c, PD_ERR_ARG, "iters must be >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = ensure_blocks(c)) return e;
  if (int e = ensure_events(c, 2)) return e;
  const pd::SpmvView A = spmv_view(c);
  const bool f32 = c->mat_prec == 1;
  const pd::MxArgs M = f32 ? mx_args(c, 1) : pd::MxArgs{};
  auto launch = [&]() {
    if (f32)
      pd::launch_spmv_f32(M, c->p32, c->y32, c->stream);
    else
      pd::launch_spmv_plain(A, c->sol, c->y, c->stream);
  };
  launch();  // warm
  CHECK_LAUNCH(c, "spmv");
  HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
  for (int i = 0; i < iters; ++i) launch();
  HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (ms_per_launch) *ms_per_launch = elapsed(c->ev[0], c->ev[1]) / iters;
  if (bytes_per_launch) pd_spmv_bytes(c, bytes_per_launch);
  return PD_OK;
}

int pd_set_specimens(pd_ctx* c, int32_t count, const int32_t* first, const int32_t* num) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  HIPCHK(c, hipSetDevice(c->device));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  reset_operator_times(c);
  const int32_t n = c->n;
  if (count == 0) {  // back to one specimen over all rows
    dfree(c->tile_spec);
    dfree(c->spec_tile0);
    dfree(c->spec_ntiles);
    c->nspec = 1;
    c->h_spec_first.clear();
    c->h_spec_count.clear();
    c->last_iters = c->last_mx_iters = 0;
    return PD_OK;
  }
  if (count < 0 || count > std::max(n, 1)) return fail(c, PD_ERR_ARG, "specimen count out of range");
  if (!first || !num) return fail(c, PD_ERR_ARG, "null specimen table");
  for (int32_t k = 0; k < count; ++k) {
    const int64_t end = (k + 1 < count) ? first[k + 1] : n;
    if (first[k] < 0 || first[k] % PD_BLOCK != 0)
      return fail(c, PD_ERR_ARG, "specimen first rows must be multiples of 256");
    if (num[k] < 1 || (int64_t)first[k] + num[k] > end)
      return fail(c, PD_ERR_ARG, "specimen rows must be non-empty, ascending and disjoint");
  }
  // rows outside every specimen must be INVISIBLE (they contribute zeros)
  std::vector<int32_t> act(n);
  if (n > 0) {
    HIPCHK(c, hipMemcpy(act.data(), c->act, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
  }
  std::vector<char> inside(n, 0);
  for (int32_t k = 0; k < count; ++k)
    for (int32_t r = first[k]; r < first[k] + num[k]; ++r) inside[r] = 1;
  for (int32_t r = 0; r < n; ++r)
    if (!inside[r] && act[r] != PD_INVISIBLE)
      return fail(c, PD_ERR_ARG, "rows outside the specimens must be INVISIBLE");
  const int32_t ntiles = (n + PD_BLOCK - 1) / PD_BLOCK;
  std::vector<int32_t> tile_spec(std::max(ntiles, 1), 0), t0(count), nt(count);
  for (int32_t k = 0; k < count; ++k) {
    t0[k] = first[k] / PD_BLOCK;
    nt[k] = (first[k] + num[k] + PD_BLOCK - 1) / PD_BLOCK - t0[k];
    const int32_t hi = (k + 1 < count) ? first[k + 1] / PD_BLOCK : ntiles;
    for (int32_t t = t0[k]; t < hi; ++t) tile_spec[t] = k;
  }
  if (count > c->state_cap) {
    dfree(c->st);
    if (int e = dalloc(c, &c->st, (size_t)count)) return e;
    if (c->h_st) (void)hipHostFree(c->h_st);
    c->h_st = nullptr;
    if (hipHostMalloc((void**)&c->h_st, sizeof(pd_cg_state) * count, hipHostMallocDefault) !=
        hipSuccess)
      return fail(c, PD_ERR_HIP, "hipHostMalloc failed");
    if (c->mx) {
      dfree(c->mx);
      if (int e = dalloc(c, &c->mx, (size_t)count)) return e;
      if (c->h_mx) (void)hipHostFree(c->h_mx);
      c->h_mx = nullptr;
      if (hipHostMalloc((void**)&c->h_mx, sizeof(pd_mx_state) * count, hipHostMallocDefault) !=
          hipSuccess)
        return fail(c, PD_ERR_HIP, "hipHostMalloc failed");
    }
    c->state_cap = count;
  }
  dfree(c->tile_spec);
  dfree(c->spec_tile0);
  dfree(c->spec_ntiles);
  if (int e = dalloc(c, &c->tile_spec, tile_spec.size())) return e;
  if (int e = dalloc(c, &c->spec_tile0, (size_t)count)) return e;
  if (int e = dalloc(c, &c->spec_ntiles, (size_t)count)) return e;
  HIPCHK(c, hipMemcpyAsync(c->tile_spec, tile_spec.data(), sizeof(int32_t) * tile_spec.size(),
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->spec_tile0, t0.data(), sizeof(int32_t) * count,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(c->spec_ntiles, nt.data(), sizeof(int32_t) * count,
                           hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->nspec = count;
  c->h_spec_first.assign(first, first + count);
  c->h_spec_count.assign(num, num + count);
  c->last_iters = c->last_mx_iters = 0;
  c->last_solve_prec = -1;
  return PD_OK;
}

int pd_get_specimen_stats(pd_ctx* c, int32_t count, pd_specimen_stats* out) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!out || count != c->nspec) return fail(c, PD_ERR_ARG, "count must equal the specimen count");
  if (c->last_solve_prec < 0) return fail(c, PD_ERR_STATE, "no solve since the specimens were set");
  for (int32_t k = 0; k < count; ++k) {
    pd_specimen_stats& o = out[k];
    if (c->last_solve_prec == 1) {
      const pd_mx_state& h = c->h_mx[k];
      o.cg_iterations = h.iter;
      o.converged = h.converged;
      o.residual_norm = h.rnorm;
      o.b_norm = h.bnorm;
      o.refinements = h.refinements;
    } else {
      const pd_cg_state& h = c->h_st[k];
      o.cg_iterations = h.iter;
      o.converged = h.converged;
      o.residual_norm = h.rnorm;
      o.b_norm = h.bnorm;
      o.refinements = 0;
    }
  }
  return PD_OK;
}

int pd_tune_spmv(pd_ctx* c, int32_t iters, double* ms3, int32_t* best) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (c->mat_prec < 0) return fail(c, PD_ERR_STATE, "no matrix assembled");
  if (c->mf) return fail(c, PD_ERR_STATE, "matrix-free operator: no stored-block SpMV variants");
  if (iters < 1) return fail(c, PD_ERR_ARG, "iters must be >= 1");
  HIPCHK(c, hipSetDevice(c->device));
  if (int e = ensure_blocks(c)) return e;
  if (int e = ensure_events(c, 2)) return e;
  const int prec = c->mat_prec;
  const int variants[3] = {2, 4, 8};
  double ms[3];
  const int keep = c->unroll[prec];
  for (int v = 0; v < 3; ++v) {
    c->unroll[prec] = variants[v];
    auto launch = [&]() {
      if (prec == 1)
        pd::launch_spmv_f32(mx_args(c, 1), c->p32, c->y32, c->stream);
      else
        pd::launch_spmv_plain(spmv_view(c), c->sol, c->y, c->stream);
    };
    launch();  // warm
    HIPCHK(c, hipEventRecord(c->ev[0], c->stream));
    for (int i = 0; i < iters; ++i) launch();
    HIPCHK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    CHECK_LAUNCH(c, "spmv variant");
    ms[v] = elapsed(c->ev[0], c->ev[1]) / iters;
  }
  int b = 0;
  for (int v = 1; v < 3; ++v)
    if (ms[v] < ms[b]) b = v;
  c->tuned[prec] = variants[b];
  c->unroll[prec] = keep;
  if (ms3)
    for (int v = 0; v < 3; ++v) ms3[v] = ms[v];
  if (best) *best = variants[b];
  return PD_OK;
}

int pd_learned_set(pd_ctx* c, const float* params, int32_t count) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!params || count != PC_NPARAM)
    return fail(c, PD_ERR_ARG, "the network takes PD_LEARNED_PARAMS floats");
  HIPCHK(c, hipSetDevice(c->device));
  if (!c->pc_p) {
    if (int e = dalloc(c, &c->pc_p, PC_NPARAM)) return e;
    if (int e = dalloc(c, &c->pc_w, PC_NWEIGHT)) return e;
  }
  HIPCHK(c, hipMemcpyAsync(c->pc_p, params, sizeof(float) * PC_NPARAM, hipMemcpyHostToDevice,
                           c->stream));
  pd::launch_pc_to_bf16(c->pc_p, c->pc_w, c->stream);
  CHECK_LAUNCH(c, "bf16 weights");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->pc_loaded = true;
  return PD_OK;
}

int pd_get_precond(pd_ctx* c, double* minv9) {
  if (!c) return fail(nullptr, PD_ERR_ARG, "null context");
  if (!minv9) return fail(c, PD_ERR_ARG, "null output");
  HIPCHK(c, hipSetDevice(c->device));
  std::vector<double> mv((size_t)c->nslices * 64 * 6);
  if (!mv.empty()) {
    if (c->last_learned) {  // the learned blocks (also after a fallback)
      HIPCHK(c, hipMemcpyAsync(mv.data(), c->pc_minv, mv.size() * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    } else if (c->mat_prec == 1) {  // the mixed assembly's float inverse
      std::vector<float> mf(mv.size());
      HIPCHK(c, hipMemcpyAsync(mf.data(), c->minv32, mf.size() * sizeof(float),
                               hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < mv.size(); ++i) mv[i] = (double)mf[i];
    } else {
      HIPCHK(c, hipMemcpyAsync(mv.data(), c->minv, mv.size() * sizeof(double),
                               hipMemcpyDeviceToHost, c->stream));
    }
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int32_t p = 0; p < c->n; ++p) {
    const double* six = mv.data() + (size_t)(p >> 6) * 384 + (p & 63);
    const double m[9] = {six[0],   six[64],  six[128], six[64], six[192],
                         six[256], six[128], six[256], six[320]};
    memcpy(minv9 + 9 * (size_t)p, m, sizeof(m));
  }
  return PD_OK;
}

// Device scratch of pd_trainer_fit_noalloc, one copy per GPU for the life of
// the process: rows of per-wave gradient partials (PC_NWEIGHT + 1 floats
// each, for batches up to PD_FIT_NOALLOC_MAX_BATCH samples), then one loss
// sum per step (up to PD_FIT_NOALLOC_MAX_STEPS).  k_pc_train writes every
// entry of a wave's row and k_pc_adam every loss it is given, so nothing
// needs clearing between calls.  Host code never passes `scratch` itself to
// a kernel (in host code the name is not a device address): it passes the
// address hipGetSymbolAddress returned for the current device.
#define PD_FIT_PARTIAL_FLOATS ((PD_FIT_NOALLOC_MAX_BATCH / 32) * (PC_NWEIGHT + 1))
#define PD_FIT_SCRATCH_FLOATS (PD_FIT_PARTIAL_FLOATS + PD_FIT_NOALLOC_MAX_STEPS)
static __device__ __align__(256) float scratch[PD_FIT_SCRATCH_FLOATS];

#define PD_SCRATCH_DEVICES 64
static std::mutex g_scratch_mu;
static float* g_scratch_ptr[PD_SCRATCH_DEVICES];

// The device address of `scratch` on `device`, looked up once per device
// (the device must be current).  Called by pd_trainer_create, so a trainer
// has its GPU's address before any fit.
static int scratch_address(int device, float** out, std::string& err) {
  if (device < 0 || device >= PD_SCRATCH_DEVICES) {
    err = "device ordinal outside the scratch table";
    return PD_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return PD_ERR_HIP;
    }
    g_scratch_ptr[device] = (float*)ptr;
  }
  *out = g_scratch_ptr[device];
  return PD_OK;
}

static int tfail(pd_trainer* t, int code, const std::string& msg) {
  if (t) t->err = msg;
  g_last_error = msg;
  return code;
}
#define THIPCHK(t, expr)                                                              \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return tfail((t), PD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

int pd_trainer_create(int32_t device, int64_t capacity, uint64_t seed, pd_trainer** out) {
  if (!out) return tfail(nullptr, PD_ERR_ARG, "null output pointer");
  *out = nullptr;
  if (capacity < 1) return tfail(nullptr, PD_ERR_ARG, "capacity must be >= 1");
  pd_trainer* t = new pd_trainer();
  t->device = device;
  t->cap = capacity;
  auto bail = [&](int code) {
    std::string msg = t->err;
    pd_trainer_destroy(t);
    g_last_error = msg;
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return bail(tfail(t, PD_ERR_ARG, "bad device"));
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(tfail(t, PD_ERR_HIP, "hipStreamCreate failed"));
  {  // the no-allocation fit's scratch on this device (allocates nothing)
    float* sp = nullptr;
    std::string msg;
    if (int e = scratch_address(device, &sp, msg)) return bail(tfail(t, e, msg));
  }
  if (hipMalloc((void**)&t->samples, sizeof(float) * PC_SAMPLE * capacity) != hipSuccess ||
      hipMalloc((void**)&t->p, sizeof(float) * PC_NPARAM) != hipSuccess ||
      hipMalloc((void**)&t->m, sizeof(float) * PC_NWEIGHT) != hipSuccess ||
      hipMalloc((void**)&t->v, sizeof(float) * PC_NWEIGHT) != hipSuccess ||
      hipMalloc((void**)&t->w, sizeof(__bf16) * PC_NWEIGHT) != hipSuccess)
    return bail(tfail(t, PD_ERR_HIP, "hipMalloc failed"));
  // He-uniform weights from a seeded splitmix64, zero biases, a small last
  // layer (S starts near I: block Jacobi), eps 1e-3
  std::vector<float> p(PC_NPARAM, 0.0f);
  uint64_t z = seed;
  auto uni = [&]() {
    z += 0x9E3779B97F4A7C15ull;
    uint64_t x = z;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (double)(x >> 11) * (1.0 / 9007199254740992.0);
  };
  auto fill = [&](int off, int rows, int cols, double scale) {
    const double lim = scale * sqrt(6.0 / cols);
    for (int i = 0; i < rows * cols; ++i) p[off + i] = (float)((2.0 * uni() - 1.0) * lim);
  };
  fill(PC_OFF_W1, PC_NH, PC_NIN, 1.0);
  fill(PC_OFF_W2, PC_NH, PC_NH, 1.0);
  fill(PC_OFF_W3, PC_NOUT, PC_NH, 0.01);
  for (int k = 0; k < PC_NIN; ++k) p[PC_OFF_ISTD + k] = 1.0f;
  p[PC_OFF_EPS] = 1e-3f;
  if (hipMemcpy(t->p, p.data(), sizeof(float) * PC_NPARAM, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(t->m, 0, sizeof(float) * PC_NWEIGHT) != hipSuccess ||
      hipMemset(t->v, 0, sizeof(float) * PC_NWEIGHT) != hipSuccess)
    return bail(tfail(t, PD_ERR_HIP, "trainer initialisation failed"));
  pd::launch_pc_to_bf16(t->p, t->w, t->stream);
  if (hipStreamSynchronize(t->stream) != hipSuccess)
    return bail(tfail(t, PD_ERR_HIP, "trainer initialisation failed"));
  *out = t;
  return PD_OK;
}

int pd_trainer_destroy(pd_trainer* t) {
  if (!t) return PD_OK;
  (void)hipSetDevice(t->device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  for (void* q : {(void*)t->samples, (void*)t->p, (void*)t->m, (void*)t->v, (void*)t->w,
                  (void*)t->partial, (void*)t->loss})
    if (q) (void)hipFree(q);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
  return PD_OK;
}

const char* pd_trainer_last_error(const pd_trainer* t) {
  return t ? t->err.c_str() : g_last_error.c_str();
}

int pd_trainer_add(pd_trainer* t, pd_ctx* c, uint64_t seed, int32_t stride, int64_t* added) {
  if (!t || !c) return tfail(t, PD_ERR_ARG, "null trainer or context");
  if (c->device != t->device) return tfail(t, PD_ERR_ARG, "trainer and context on different devices");
  if (c->mat_prec < 0) return tfail(t, PD_ERR_STATE, "no matrix assembled in the context");
  if (stride < 1) return tfail(t, PD_ERR_ARG, "stride must be >= 1");
  THIPCHK(t, hipSetDevice(c->device));
  if (ensure_blocks(c) != PD_OK) return tfail(t, PD_ERR_HIP, "writing the deferred blocks failed");
  const int32_t n = c->n;
  uint64_t h = seed;
  h = (h ^ (h >> 30)) * 0xBF58476D1CE4E5B9ull;
  h = (h ^ (h >> 27)) * 0x94D049BB133111EBull;
  h ^= h >> 31;
  const int64_t off = (int64_t)(h % (uint64_t)stride);
  const int64_t rows = off < n ? (n - 1 - off) / stride + 1 : 0;
  if (t->count + rows > t->cap) return tfail(t, PD_ERR_ARG, "trainer capacity exceeded");
  if (!c->pc_feat)
    if (int e = dalloc(c, &c->pc_feat, (size_t)std::max(n, 1) * PC_NIN))
      return tfail(t, e, c->err);
  pd::PcFeatArgs fa;
  fa.n = n;
  fa.slice_off = c->slice_off;
  fa.top = c->top;
  fa.cap = c->cap;
  fa.col = c->col;
  fa.x = c->x;
  fa.diag = c->diag;
  fa.bond_stats = c->bond_stats;
  fa.act = c->act;
  fa.body = c->body;
  fa.mat = c->dmat;
  fa.feat = c->pc_feat;
  pd::launch_pc_features(fa, c->stream);
  pd::launch_pc_probe(n, seed, c->z, c->stream);
  if (c->mat_prec == 1)
    pd::launch_spmv_wide(mx_args(c, 1), c->z, c->y, c->stream);
  else
    pd::launch_spmv_plain(spmv_view(c), c->z, c->y, c->stream);
  pd::PcSampleArgs sa;
  sa.n = n;
  sa.feat = c->pc_feat;
  sa.diag = c->diag;
  sa.Ae = c->y;
  sa.e = c->z;
  sa.act = c->act;
  sa.bc = c->bc;
  sa.seed = seed;
  sa.keep = (uint32_t)stride;
  sa.out = t->samples + t->count * PC_SAMPLE;
  sa.capacity = t->cap - t->count;
  pd::launch_pc_samples(sa, c->stream);
  THIPCHK(t, hipGetLastError());
  THIPCHK(t, hipStreamSynchronize(c->stream));
  t->count += rows;
  if (added) *added = rows;
  return PD_OK;
}

}  // extern "C"

namespace {
// The fit loop of pd_trainer_fit and pd_trainer_fit_noalloc: `steps` Adam
// steps on batches drawn from the trainer's samples, per-wave gradient
// partials in `partial` (batch / 32 rows), loss sums in `loss` (steps).
int run_fit(pd_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed, float* partial,
            float* loss, float* losses) {
  const int32_t waves = batch / 32;
  if (!t->normed) {
    pd::launch_pc_norm(t->samples, t->count, t->p, t->stream);
    t->normed = true;
  }
  for (int32_t k = 0; k < steps; ++k) {
    pd::PcTrainArgs a;
    a.samples = t->samples;
    a.nsamples = t->count;
    a.batch = batch;
    a.seed = seed;
    a.step = t->steps_done;
    a.w = t->w;
    a.p = t->p;
    a.partial = partial;
    pd::launch_pc_train(a, t->stream);
    pd::launch_pc_adam(partial, waves, t->p, t->m, t->v, t->w, t->steps_done, lr, loss + k,
                       t->stream);
    ++t->steps_done;
  }
  THIPCHK(t, hipGetLastError());
  if (losses)
    THIPCHK(t, hipMemcpyAsync(losses, loss, sizeof(float) * steps, hipMemcpyDeviceToHost,
                              t->stream));
  THIPCHK(t, hipStreamSynchronize(t->stream));
  if (losses)
    for (int32_t k = 0; k < steps; ++k) losses[k] /= (float)batch;  // mean |M v - e|^2
  return PD_OK;
}

int check_fit(pd_trainer* t, int32_t steps, int32_t batch) {
  if (!t) return tfail(nullptr, PD_ERR_ARG, "null trainer");
  if (t->count < 1) return tfail(t, PD_ERR_STATE, "no samples");
  if (steps < 1 || batch < 128 || batch % 128 != 0)
    return tfail(t, PD_ERR_ARG, "steps >= 1 and batch a positive multiple of 128");
  return PD_OK;
}
}  // namespace

extern "C" {

int pd_trainer_fit(pd_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed,
                   float* losses) {
  if (int e = check_fit(t, steps, batch)) return e;
  THIPCHK(t, hipSetDevice(t->device));
  const int32_t waves = batch / 32;
  if (waves > t->partial_waves) {
    if (t->partial) (void)hipFree(t->partial);
    t->partial = nullptr;
    THIPCHK(t, hipMalloc((void**)&t->partial, sizeof(float) * (PC_NWEIGHT + 1) * waves));
    t->partial_waves = waves;
  }
  if (steps > t->loss_cap) {
    if (t->loss) (void)hipFree(t->loss);
    t->loss = nullptr;
    THIPCHK(t, hipMalloc((void**)&t->loss, sizeof(float) * steps));
    t->loss_cap = steps;
  }
  return run_fit(t, steps, batch, lr, seed, t->partial, t->loss, losses);
}

int pd_trainer_fit_noalloc(pd_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed,
                           float* losses) {
  if (int e = check_fit(t, steps, batch)) return e;
  if (batch > PD_FIT_NOALLOC_MAX_BATCH)
    return tfail(t, PD_ERR_ARG,
                 "batch " + std::to_string(batch) + " exceeds the no-allocation fit's maximum " +
                     std::to_string(PD_FIT_NOALLOC_MAX_BATCH));
  if (steps > PD_FIT_NOALLOC_MAX_STEPS)
    return tfail(t, PD_ERR_ARG,
                 "steps " + std::to_string(steps) + " exceeds the no-allocation fit's maximum " +
                     std::to_string(PD_FIT_NOALLOC_MAX_STEPS));
  THIPCHK(t, hipSetDevice(t->device));
  float* sp = nullptr;
  std::string msg;
  if (int e = scratch_address(t->device, &sp, msg)) return tfail(t, e, msg);
  return run_fit(t, steps, batch, lr, seed, sp, sp + PD_FIT_PARTIAL_FLOATS, losses);
}

int pd_trainer_params(pd_trainer* t, float* params, int32_t count) {
  if (!t || !params || count != PC_NPARAM) return tfail(t, PD_ERR_ARG, "bad arguments");
  THIPCHK(t, hipSetDevice(t->device));
  THIPCHK(t, hipMemcpyAsync(params, t->p, sizeof(float) * PC_NPARAM, hipMemcpyDeviceToHost,
                            t->stream));
  THIPCHK(t, hipStreamSynchronize(t->stream));
  return PD_OK;
}

int pd_trainer_samples(pd_trainer* t, int64_t first, int64_t count, float* out) {
  if (!t || !out || first < 0 || count < 0 || first + count > t->count)
    return tfail(t, PD_ERR_ARG, "bad sample range");
  THIPCHK(t, hipSetDevice(t->device));
  if (count)
    THIPCHK(t, hipMemcpy(out, t->samples + first * PC_SAMPLE, sizeof(float) * PC_SAMPLE * count,
                         hipMemcpyDeviceToHost));
  return PD_OK;
}

}  // extern "C"
