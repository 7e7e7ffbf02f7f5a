/*
 * pd_solver.h -- C ABI of the MI355X implicit bond-based peridynamics
 * engine.  Plain C types only; the library owns device memory and one HIP
 * stream per context.  Every call returns a pd_status; pd_last_error()
 * returns the message of the most recent failure.
 *
 * Each entry point replaces a launch site or host routine of the reference
 * (SCIInstitute/SCI-Solver_Peridynamic, src/); the replaced file:line is
 * given on each declaration.
 *
 * Units: whatever the caller uses consistently (the reference's scaled
 * units or SI); the engine adds no scale factor.
 * Vectors: real4 layout, 4 doubles per particle with the w lane ignored
 * on input and written as 0 (core/definitions.h:37-39).
 */
#ifndef PD_SOLVER_H
#define PD_SOLVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PD_ABI_VERSION 1
#define PD_MAX_BODIES 8
#define PD_MAX_BONDS_LIMIT 1024 /* upper bound accepted for max_bonds */

typedef enum pd_status {
  PD_OK = 0,
  PD_ERR_HIP = 1,            /* HIP runtime failure (message has details) */
  PD_ERR_ARG = 2,            /* invalid argument or size mismatch */
  PD_ERR_NOT_CONVERGED = 3,  /* CG reached max_iter before rtol */
  PD_ERR_BOND_OVERFLOW = 4,  /* a particle has more than max_bonds bonds */
  PD_ERR_STATE = 5,          /* call out of order (e.g. step before bonds) */
  PD_ERR_NO_DEVICE = 6       /* no usable HIP device */
} pd_status;

/* Material of one body (the reference has one body; these are its
 * pd_horizon, pd_particle_radius, pd_C_times_V*PD_K and pd_particle_mass,
 * core/parameters.cu:351-370). */
typedef struct pd_material {
  double horizon; /* bond exists iff dist < horizon (strict) */
  double radius;  /* volume-correction width r (cg_solver.cuh:281-284) */
  double c_bond;  /* force per unit stretch, C*K */
  double mass;    /* particle mass */
} pd_material;

typedef struct pd_config {
  int32_t num_particles;
  int32_t device;     /* HIP device ordinal (reference gpu_device) */
  int32_t max_bonds;  /* per-particle cap, 528 in the reference */
  int32_t num_bodies; /* 1..PD_MAX_BODIES */
  pd_material material[PD_MAX_BODIES];
} pd_config;

/* Fields for pd_upload / pd_download */
typedef enum pd_field {
  PD_FIELD_POSITION = 0,          /* double4[N]  pd_position */
  PD_FIELD_VELOCITY = 1,          /* double4[N]  pd_velocity */
  PD_FIELD_ORIGINAL_POSITION = 2, /* double4[N]  pd_original_pos */
  PD_FIELD_STRETCH = 3,           /* double[N]   pd_stretch */
  PD_FIELD_ORIGINAL_STRETCH = 4,  /* double[N]   pd_original_stretch */
  PD_FIELD_ACTIVITY = 5,          /* int32[N]    pd_activity */
  PD_FIELD_BODY = 6,              /* int32[N]    body index */
  PD_FIELD_BC_FLAG = 7,           /* int32[N]    1 = prescribed velocity */
  PD_FIELD_BC_VELOCITY = 8,       /* double4[N]  prescribed velocity */
  PD_FIELD_SOLUTION = 9,          /* double4[N]  pd_system_solution */
  PD_FIELD_FORCE = 10,            /* double4[N]  pd_force (download) */
  PD_FIELD_RHS = 11,              /* double4[N]  pd_system_vector */
  PD_FIELD_NEW_STRETCH = 12,      /* double[N]   pd_new_stretch */
  PD_FIELD_BOND_TOP = 13,         /* int32[N]    pd_bond_list_top (upload:
                                   *              -1 <= top < capacity) */
  PD_FIELD_CONTACT_FORCE = 14,    /* double4[N]  contact force */
  PD_FIELD_COUNT = 15
} pd_field;

typedef struct pd_step_params {
  double dt;          /* seconds; reference: pd_time_step * 2^-23 */
  double clock_scale; /* phase-in factor (simulator.cu:1163-1167) */
  double gravity;     /* 9.8 in the reference, applied along -y */
  int32_t integrator; /* 0 implicit Euler, 1 Newmark-beta */
  int32_t precond;    /* 0 identity (reference CUSP), 1 block Jacobi,
                         2 learned blocks (pd_learned_set) */
  double rtol;        /* stop when ||r|| <= rtol * ||S b|| */
  int32_t max_iter;   /* CG iteration limit (reference 100000) */
  int32_t walls;      /* 1: clamp positions to [bmin+r, bmax-r] */
  double bmin[3];
  double bmax[3];
  double wall_radius;
  double restitution;       /* 0.001 in the reference */
  double contact_radius;    /* Rc; 0 disables contact */
  double contact_stiffness; /* Kc */
  double contact_volume;    /* PD_V divisor */
  double contact_cell;      /* grid cell edge for contact, >= Rc */
  int32_t track_gap;        /* 1: report the bond closest to breaking */
  int32_t timing;           /* 1: time phases and SpMV launches */
  /* Solver precision (not in the reference, which is FP64 only).
   * 0: FP64 CG.  1: the assembly also rounds the bond blocks, the diagonal
   * and its block-Jacobi inverse to float, and the solve is iterative
   * refinement: r = S(b - A x) of the FP64 system in FP64 (its bond blocks
   * recomputed), stop when ||r|| <= rtol ||S b||, else a PCG on float
   * vectors solves A32 d = r until ||r32|| <= max(inner_rtol ||r||, tol / 2)
   * and x += d in FP64.  Dot products and norms are FP64 throughout. */
  int32_t precision;
  int32_t max_refine;     /* precision 1: refinement limit (0: 20) */
  double inner_rtol;      /* precision 1: inner reduction (0: 1e-5) */
  int32_t inner_max_iter; /* precision 1: per refinement (0: max_iter) */
  /* contact pairs: 0 particles of different bodies; 1 every pair closer
   * than contact_radius, same body included (the pair set of the
   * reference's collidePDParticles, core/simulator.cuh:672-765) */
  int32_t contact_mode;
  /* SpMV slots per trip (2, 4 or 8; the summation order, and so every
   * result, is the same for all); 0: the value pd_tune_spmv picked for this
   * precision, else 4 */
  int32_t spmv_unroll;
  /* pd_step only: 1 replays the step from hipGraphs captured on first use
   * (assembly, stretch commit, the solve's start and as many iterations as
   * the last step needed, then the position update), one launch and one
   * wait per step; the host-driven loop finishes any specimen still
   * iterating.  Needs contact off and track_gap 0.  Results are identical
   * to graph = 0. */
  int32_t graph;
  /* precond 2: the learned solve gets at most learned_fallback x the
   * iterations of this context's last block-Jacobi solve (0: 1.5), then
   * continues with block Jacobi from its iterate */
  double learned_fallback;
  /* Operator of the FP64 solve (not in the reference, which stores A).
   * 0: the assembly stores every bond block (3x3 BSR) and the SpMV streams
   * them.  1: matrix-free: the assembly stores the diagonal, b and the
   * block-Jacobi inverse only, and every SpMV recomputes each live bond's
   * block from the step's positions with the assembly's own arithmetic,
   * so every product, and every result, is identical to 0.  Needs
   * precision 0; pd_get_matrix then fails (no stored bond blocks), and
   * pd_spmv, the learned preconditioner's trainer and every solve apply
   * the recomputed operator.  2 (pd_step; pd_assemble takes the current
   * choice): per step, whichever of 0 and 1 is cheaper, from this
   * context's own timings of both products and the previous step's
   * iteration count (the first steps time each operator in turn, with the
   * timing of their products on); precision 1 streams its float blocks. */
  int32_t matrix_free;
  /* precond 2, pd_step only: the fallback's baseline is the iteration
   * count of this context's latest block-Jacobi solve.  A precond-2 step is
   * a probe when the context has no such count yet, when the previous
   * precond-2 step fell back to block Jacobi, or when learned_probe
   * precond-2 steps have run since the count was refreshed (0: 16;
   * negative: no periodic probes): its solve uses the block-Jacobi inverse
   * instead of the learned blocks and refreshes the count.  pd_solve never
   * probes. */
  int32_t learned_probe;
} pd_step_params;

typedef struct pd_step_stats {
  int32_t cg_iterations;
  int32_t converged;
  double residual_norm;
  double b_norm;
  int64_t broken_bonds; /* directed bonds removed this step */
  int64_t live_bonds;   /* directed bonds alive after the step */
  double min_gap;       /* min |stretch - threshold| (track_gap) */
  int32_t gap_p;
  int32_t gap_q;
  /* filled when timing = 1 (HIP events on the context stream) */
  double ms_contact;
  double ms_assemble;
  double ms_solve;
  double ms_update;
  double ms_spmv_total; /* sum over the CG SpMV launches that ran, streaming
                           stored blocks or recomputing them (matrix-free) */
  int32_t spmv_launches;
  int32_t host_syncs;
  int32_t refinements;    /* precision 1: FP64 corrections applied */
  double ms_resid_total;  /* precision 1: FP64 residual launches that ran */
  int32_t resid_launches;
  int32_t precond_fallbacks; /* precond 2: 1 if block Jacobi took over */
  /* stored-block FP64 steps: the solve's first product, which also writes
   * the bond blocks the assembly left to it; mixed steps: the first inner
   * product, which writes the float blocks (neither in ms_spmv_total) */
  double ms_spmv_store;
  int32_t precond_probe; /* precond 2: 1 if this step's solve was a
                            block-Jacobi probe (learned_probe) */
  int32_t operator_mf;   /* 1 if this step's products recomputed the bond
                            blocks (matrix_free 1, or 2's choice) */
  int32_t contact_grid;  /* the contact grid build of this step: 0 none
                            (contact off or no candidates), 1 / 2 the
                            one-workgroup build with one / two radix passes,
                            3 the multi-launch build */
} pd_step_stats;

/* Solver result of one specimen (pd_set_specimens). */
typedef struct pd_specimen_stats {
  int32_t cg_iterations;
  int32_t converged;
  double residual_norm;
  double b_norm;
  int32_t refinements;
} pd_specimen_stats;

typedef struct pd_bond_stats {
  int64_t total_bonds; /* directed bonds */
  int32_t min_bonds;
  int32_t max_bonds;
  int64_t slots;       /* padded slot count of the sliced layout */
  double ms_build;
} pd_bond_stats;

typedef struct pd_ctx pd_ctx;

/* Simulator::Simulator + MemoryManager::MemoryManager
 * (core/simulator.cu:34-54, core/memory_manager.cu:16-217): select
 * cfg->device, allocate every per-particle buffer, create the stream. */
int pd_ctx_create(const pd_config* cfg, pd_ctx** out);

/* Simulator::~Simulator / MemoryManager::~MemoryManager
 * (core/memory_manager.cu:219-233). */
int pd_ctx_destroy(pd_ctx* ctx);

/* checkCudaErrors / TRUE_OR_DIE messages (core/helper_cuda.h:892-922,
 * core/utilities.h:27-35), returned instead of exiting.  ctx may be NULL
 * for failures of pd_ctx_create. */
const char* pd_last_error(const pd_ctx* ctx);

/* MemoryManager::uploadToDevice (core/memory_manager.cu:236-260):
 * synchronous host-to-device copy of one field; bytes must equal the
 * field size for num_particles. */
int pd_upload(pd_ctx* ctx, int32_t field, const void* src, size_t bytes);

/* MemoryManager::downloadFromDevice (core/memory_manager.cu:274-298),
 * without the scale factor (the caller owns units). */
int pd_download(pd_ctx* ctx, int32_t field, void* dst, size_t bytes);

/* Simulator::initPeridynamicsBonds bond part (core/simulator.cu:825-943,
 * core/kdtree.h:141-353): GPU uniform-grid hash of the current positions,
 * LDS radix sort, cell start/end, then for every particle the sorted list
 * of q != p in the same body with dist(p,q) < horizon[body].  Sets
 * pd_bond_list_top = count - 1.  Fails with PD_ERR_BOND_OVERFLOW if a
 * particle exceeds max_bonds (the reference silently truncated). */
int pd_build_bonds(pd_ctx* ctx, pd_bond_stats* stats);

/* Load a host-built bond list instead (e.g. a restored state):
 * counts[N], lists row-major with row stride `stride`, each row sorted. */
int pd_set_bonds(pd_ctx* ctx, const int32_t* counts, const int32_t* lists,
                 int32_t stride);

/* Current bond lists in list order (live bonds first, then the broken ones
 * that swap-remove moved past the top): capacity[N] and lists row-major
 * with row stride `stride` >= max capacity. */
int pd_get_bonds(pd_ctx* ctx, int32_t* capacity, int32_t* lists,
                 int32_t stride);

/* Simulator::printDebugInfoLinearSystem (core/simulator.cu:656-732) as
 * data: the assembled system in the reference's row-major Mat3x3 layout,
 * diag9[N*9] = A_pp and vals9[(p*stride + k)*9] = A_{p,list[p][k]} for
 * the live slots k <= top[p] (the remaining entries are left untouched). */
int pd_get_matrix(pd_ctx* ctx, double* diag9, double* vals9, int32_t stride);
/* (After an assembly with precision 1, vals9 holds the float blocks
 * widened to double; diag9 is the FP64 diagonal.) */

/* Simulator::integratePeridynamicsImplicitEuler (core/simulator.cu:
 * 297-467) and integratePeridynamicsNewmarkBeta (core/simulator.cu:
 * 470-652) without the adaptive rollback branch: contact (if enabled),
 * fused calculatePDForceDerivative + fill (core/cg_solver.cuh:200,
 * core/implicit_euler.cuh:19,62), stretch update (core/simulator.cuh:
 * 2601), CG solve (core/simulator.cu:735), velocity / position update
 * (core/implicit_euler.cuh:90, core/newmark_beta.cuh:89,113).
 * Returns PD_ERR_NOT_CONVERGED (state already advanced) when CG hits
 * max_iter, matching CG_convergent_ = false (core/simulator.cu:780-790). */
int pd_step(pd_ctx* ctx, const pd_step_params* prm, pd_step_stats* stats);

/* Test hook, first half of pd_step: contact + force / Jacobian / breaking
 * + fill of A and b (launch sites core/simulator.cu:306, 430, 438).  Leaves
 * the stretch limits unswapped and does not solve. */
int pd_assemble(pd_ctx* ctx, const pd_step_params* prm, pd_step_stats* stats);

/* updatePDStretchNeighborList (core/simulator.cuh:2601-2620): s <- s_new */
int pd_commit_stretch(pd_ctx* ctx);

/* PeridynamicsMatrix::operator() / computeAx (core/cg_solver.cuh:343-428):
 * y = A x for host real4 vectors x, y of num_particles entries.  After an
 * assembly with precision 1 the operator is A32 in FP64 arithmetic (the
 * refinement's residual operator). */
int pd_spmv(pd_ctx* ctx, const double* x4, double* y4);

/* Simulator::solveLinearSystem (core/simulator.cu:735-799): CG on the
 * assembled A and b, warm-started from the current solution. */
int pd_solve(pd_ctx* ctx, const pd_step_params* prm, pd_step_stats* stats);

/* MemoryManager::transferData(PD_VELOCITY, PD_SYSTEM_SOLUTION) +
 * updatePDPosition* (core/memory_manager.cu:541, core/implicit_euler.cuh:
 * 90-176, core/newmark_beta.cuh:89-199). */
int pd_update(pd_ctx* ctx, const pd_step_params* prm);

/* MemoryManager::backup{BondListTopIndex,PDPosition,PDVelocity} /
 * restore* (core/memory_manager.cu:491-538) for adaptive-dt rollback. */
int pd_backup(pd_ctx* ctx);
int pd_restore(pd_ctx* ctx);

/* Wait for all work on the context stream. */
int pd_synchronize(pd_ctx* ctx);

/* Times `iters` launches of the plain SpMV kernel y = A x on the resident
 * matrix with HIP events on the context stream; returns the mean kernel
 * time in ms and the algorithmic bytes one launch moves.  After an
 * assembly with precision 1 the kernel is the float inner SpMV. */
int pd_bench_spmv(pd_ctx* ctx, int32_t iters, double* ms_per_launch,
                  double* bytes_per_launch);

/* Self-check (no counterpart in the reference) of the shared-reciprocal
 * quotients and the unscaled square root that the assembly and the
 * matrix-free operator use when the step's operands lie in range: on
 * `samples` random operands spanning those ranges (a coordinate difference
 * over a float bond length; a current over a reference length and back; a
 * bond stiffness over a reference length; half of a horizon less a
 * reference length over a radius; the root of a squared bond length),
 * counts in *mismatches the results whose bits differ from IEEE division
 * and sqrt on this device.  0 is the expected answer. */
int pd_check_division(pd_ctx* ctx, int64_t samples, uint64_t seed, int64_t* mismatches);

/* Algorithmic HBM bytes of one CG SpMV launch on the current bonds:
 * FP64, live blocks * (6*8 + 4) + N * (6*8 + 4 + 2*24); after an assembly
 * with precision 1 (float inner SpMV), live blocks * (6*4 + 4) +
 * N * (6*4 + 4 + 2*12); after a matrix-free assembly, live bonds * (4 + 8)
 * + N * (6*8 + 4 + 24 + 4 + 2*24) (columns and reference lengths;
 * diagonal, top, positions, body, x and y: the blocks are recomputed). */
int pd_spmv_bytes(pd_ctx* ctx, double* bytes);

/* Autotuner (no counterpart in the reference): times `iters` launches of
 * each SpMV variant (2, 4 and 8 slots per trip) on the resident matrix of
 * the last assembly's precision, remembers the fastest for that precision
 * (used when spmv_unroll = 0) and returns the mean times in ms3[3] and the
 * choice in *best.  Fails after a matrix-free assembly (no stored blocks). */
int pd_tune_spmv(pd_ctx* ctx, int32_t iters, double* ms3, int32_t* best);

/* Learned block preconditioner (no counterpart in the reference).  The
 * network maps per-particle features of the assembled system (live bond
 * count, mean and max bond stretch, the diagonal block relative to m I,
 * distance to the nearest broken bond) through 16-64-64-8 ReLU layers (BF16
 * weights, FP32 accumulation) to L_p, and precond 2 uses
 * M_p = C_p^{-T} (L_p L_p^T + eps I) C_p^{-1} with D_p = C_p C_p^T in place
 * of the block-Jacobi inverse D_p^{-1} (which is the case L_p = I).
 * params: PD_LEARNED_PARAMS floats, W1[64][16] b1[64] W2[64][64] b2[64]
 * W3[8][64] b3[8] feature_mean[16] feature_inv_std[16] eps. */
#define PD_LEARNED_PARAMS 5801
int pd_learned_set(pd_ctx* ctx, const float* params, int32_t count);
/* The preconditioner blocks of the last solve, 9 doubles per particle: the
 * learned blocks M_p when it was a precond-2 solve (also one that fell back
 * to block Jacobi), else the block-Jacobi inverse (after a mixed-precision
 * assembly its float copy, widened). */
int pd_get_precond(pd_ctx* ctx, double* minv9);

/* Training of the network on systems the simulator assembled: samples are
 * (features, D_p, (A e)_p, e_p) for random sign probes e on every
 * stride-th row (offset from the seed); fitting minimises the mean of
 * |M_p (A e)_p - e_p|^2 with Adam, `steps` steps of `batch` samples drawn
 * with replacement (batch a multiple of 128).  losses[k] (if non-null)
 * receives each step's mean loss. */
typedef struct pd_trainer pd_trainer;
int pd_trainer_create(int32_t device, int64_t capacity, uint64_t seed, pd_trainer** out);
int pd_trainer_destroy(pd_trainer* t);
const char* pd_trainer_last_error(const pd_trainer* t);
int pd_trainer_add(pd_trainer* t, pd_ctx* ctx, uint64_t seed, int32_t stride, int64_t* added);
int pd_trainer_fit(pd_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed,
                   float* losses);
int pd_trainer_params(pd_trainer* t, float* params, int32_t count);

/* pd_trainer_fit without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory while they
 * run (code under HIP graph capture, tight serving or training loops,
 * harnesses that forbid hipMalloc).  pd_trainer_fit allocates its scratch
 * (per-wave gradient partials, batch / 32 rows of PD_LEARNED_PARAMS - 32
 * floats, and one loss sum per step) when a call needs more than an earlier
 * one; this variant keeps both in one library-owned device array per GPU
 * instead, sized for batch <= PD_FIT_NOALLOC_MAX_BATCH and
 * steps <= PD_FIT_NOALLOC_MAX_STEPS (larger requests fail with
 * PD_ERR_ARG before any work), and otherwise runs the same kernels with the
 * same results.  Every call on one GPU shares that array: calls that overlap
 * on one GPU (two trainers on different streams or threads) race and must
 * be serialised by the caller. */
#define PD_FIT_NOALLOC_MAX_BATCH 4096
#define PD_FIT_NOALLOC_MAX_STEPS 16384
int pd_trainer_fit_noalloc(pd_trainer* t, int32_t steps, int32_t batch, float lr, uint64_t seed,
                           float* losses);
/* samples [first, first + count), 28 floats each: features[16], D (6),
 * (A e)_p (3), e_p (3) */
int pd_trainer_samples(pd_trainer* t, int64_t first, int64_t count, float* out);

/* Independent specimens in one context (no counterpart in the reference,
 * which runs one body per process).  Specimen k owns rows [first[k],
 * first[k] + num[k]); first[k] must be a multiple of 256 and the ranges
 * ascending and disjoint; rows outside every specimen must be INVISIBLE
 * and far from the others' horizons (they carry no bonds and contribute
 * zeros).  Afterwards every solve keeps its scalars, sums, stopping test
 * and iteration count per specimen, in the same launches for all of them,
 * so each specimen's result equals a context that holds it alone.
 * count = 0 returns to one specimen over all rows. */
int pd_set_specimens(pd_ctx* ctx, int32_t count, const int32_t* first, const int32_t* num);

/* Per-specimen results of the last solve (count = the specimen count).
 * pd_step_stats then reports the most iterations, converged = all, and the
 * residual of the specimen farthest from its tolerance. */
int pd_get_specimen_stats(pd_ctx* ctx, int32_t count, pd_specimen_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* PD_SOLVER_H */
