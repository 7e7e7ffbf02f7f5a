/*
 * pd_oracle.h -- CPU restatement of the PeriCUDA implicit bond-based
 * peridynamics step (reference: SCIInstitute/SCI-Solver_Peridynamic).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never
 * links, calls or falls back to it.
 *
 * Every routine cites the reference file:line it restates.  Floating-point
 * expressions follow the reference's host compile exactly: IEEE double,
 * no FMA contraction (-ffp-contract=off), the two float roundings that the
 * reference's fmaxf calls introduce, and sums in the reference's order.
 * Where the reference delegates to third-party code (CUSP CG, Thrust
 * reductions), the order of operations is the deterministic tree written
 * down in pd_oracle.c (orc_dot) and shared with the HIP kernels.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_rand_*          glibc rand() via ctypes, first 100000 outputs
 *   orc_getstretch      closed-form mean/variance + libc replay
 *   orc_bonds_kdtree    brute force (orc_bonds_brute) on jittered lattices
 *   orc_assemble        two-particle closed form, finite differences,
 *                       translation invariance, bitwise symmetry,
 *                       rotation equivariance, breaking threshold
 *   orc_spmv            dense numpy product
 *   orc_cg              numpy dense solve, n-eigenvalue termination,
 *                       block-diagonal one-iteration termination
 *   orc_update          free-fall closed form, wall-clamp closed form
 *   orc_contact         two-particle closed form, symmetry
 *   orc_dot             exact rational sum on integer-valued data
 *   orc_cg_mixed        numpy dense solves of the FP64 system (the refined
 *                       solution) and of the float-rounded one (which it
 *                       must not reach), FP64 stopping rule on the true
 *                       residual, the float accuracy floor of one inner
 *                       solve
 */
#ifndef PD_ORACLE_H
#define PD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_BODIES 8
#define ORC_INVISIBLE 3 /* core/parameters.h:137-144 enum Activity */

/* ---- glibc TYPE_3 additive-feedback generator (rand()/srand()) ---- */
typedef struct orc_rand {
  uint32_t ring[34];
  int64_t i; /* index of the next value r_i to produce */
} orc_rand;

void orc_rand_seed(orc_rand* st, uint32_t seed);
int32_t orc_rand_next(orc_rand* st);
/* core/simulator.cu:802-823 frand()/getstretch() */
double orc_getstretch(orc_rand* st, float s0);
/* core/simulator.cu:857-864: srand(seed); s00[i] = s[i] = getstretch(s0) */
void orc_stretch_init(int32_t n, float s0, uint32_t seed, double* s00);

/* ---- bond (horizon neighbour) lists ---- */
/* core/kdtree.h:18-353 + core/simulator.cu:825-943.  Lists are written
 * row-major lists[p*max_bonds + k], sorted ascending per row (the
 * reference's std::sort at core/simulator.cu:920).  Returns the number of
 * bonds dropped by the max_bonds cap (reference add_bond, kdtree.h:338). */
int64_t orc_bonds_kdtree(int32_t n, const double* x4, double horizon,
                         int32_t max_bonds, int32_t leaf_size,
                         int32_t* counts, int32_t* lists);
/* All q != p in the same body with dist(p,q) < horizon[body[p]], the
 * reference's distance expression (kdtree.h:66-79).  Returns -1 if a row
 * exceeds max_bonds, else the total number of directed bonds. */
int64_t orc_bonds_brute(int32_t n, const double* x4, const int32_t* body,
                        const double* horizon, int32_t max_bonds,
                        int32_t* counts, int32_t* lists);

/* ---- per-step system ---- */
typedef struct orc_material {
  double horizon; /* delta: bonds need dist < horizon */
  double radius;  /* particle radius r: volume correction width */
  double c_bond;  /* C*K (pd_C_times_V * PD_K): force per unit stretch */
  double mass;    /* particle mass */
} orc_material;

typedef struct orc_system {
  int32_t n;
  int32_t nbody;
  orc_material mat[ORC_MAX_BODIES];
  /* particle arrays; real4 = 4 doubles per particle, w lane kept 0 */
  double* x;   /* position */
  double* v;   /* velocity */
  double* x0;  /* original position */
  double* sol; /* CG solution, persists as the next warm start */
  double* f;   /* force */
  double* b;   /* right-hand side */
  double* fc;  /* contact force of this step */
  double* vbc; /* prescribed velocity of Dirichlet particles */
  double* s;     /* stretch limit in use this step */
  double* s_new; /* stretch limit for the next step */
  double* s00;   /* original stretch limit */
  int32_t* act;  /* activity (INVISIBLE rows are skipped) */
  int32_t* body; /* body index into mat[] */
  int32_t* bc;   /* 1 = prescribed velocity (Dirichlet), 0 = free */
  /* bonds: row p owns slots [row_start[p], row_start[p+1]); the first
   * top[p]+1 slots are live (reference pd_bond_list_top semantics) */
  int64_t* row_start;
  int32_t* col;
  int32_t* top;
  double* val;  /* 9 doubles per slot: A_{p,col} row-major 3x3 */
  double* diag; /* 9 doubles per row: A_{pp} */
  double* minv; /* 9 doubles per row: inverse of A_{pp} (block Jacobi) */
  /* mixed precision (precision = 1): the same blocks rounded to float */
  float* val32;  /* 9 per slot */
  float* diag32; /* 9 per row */
  float* minv32; /* 9 per row: float(inverse of the FP64 diagonal block) */
} orc_system;

typedef struct orc_step_params {
  double dt;          /* seconds (reference: pd_time_step * 2^-23) */
  double clock_scale; /* phase-in factor c (core/simulator.cu:1163-1167) */
  double gravity;     /* 9.8 in the reference, along -y */
  int32_t integrator; /* 0 implicit Euler, 1 Newmark-beta */
  int32_t precond;    /* 0 identity (reference), 1 block Jacobi, 2 minv as given */
  double rtol;        /* relative tolerance on ||r|| / ||S b|| */
  int32_t max_iter;
  int32_t walls;      /* 1: clamp to [bmin+r, bmax-r] with restitution */
  double bmin[3];
  double bmax[3];
  double wall_radius;
  double restitution;       /* 0.001 in the reference */
  double contact_radius;    /* Rc; 0 disables contact */
  double contact_stiffness; /* Kc */
  double contact_volume;    /* PD_V divisor */
  double contact_cell;      /* uniform-grid cell edge, >= contact_radius */
  int32_t precision;        /* 0 FP64 CG; 1 FP32 inner CG + FP64 refinement */
  double inner_rtol;        /* inner stop: ||r32|| <= max(inner_rtol ||r||, tol/2) */
  int32_t inner_max_iter;   /* per refinement */
  int32_t max_refine;       /* outer FP64 refinements */
  int32_t contact_mode;     /* 0 pairs of different bodies, 1 all pairs */
} orc_step_params;

typedef struct orc_step_stats {
  int32_t cg_iterations;
  int32_t converged;
  double residual_norm;
  double b_norm;
  int64_t broken_bonds; /* directed bonds removed this step */
  int64_t live_bonds;   /* directed bonds alive after this step */
  double min_gap;       /* min |sigma - threshold| over bonds tested */
  int32_t gap_p;        /* row of that bond */
  int32_t gap_q;        /* column of that bond */
  int32_t refinements;  /* mixed precision: FP64 corrections applied */
} orc_step_stats;

/* Deterministic reduction shared with the HIP kernels (see pd_oracle.c). */
double orc_dot(int32_t n, const double* a4, const double* b4);

/* core/cg_solver.cuh:200-339 fused with core/implicit_euler.cuh:19-86
 * (integrator 0) or core/newmark_beta.cuh:19-86 (integrator 1). */
void orc_assemble(orc_system* sys, const orc_step_params* prm,
                  orc_step_stats* st);
/* core/cg_solver.cuh:343-383 computeAx: y = A x on real4 vectors */
void orc_spmv(const orc_system* sys, const double* x4, double* y4);
/* block-Jacobi inverse of every diagonal block (new; the reference's
 * scalar diagonal preconditioner is commented out, simulator.cu:768) */
void orc_block_jacobi(orc_system* sys);
/* core/simulator.cu:735-799 (CUSP krylov::cg structure), projected onto
 * the free rows when bc[] marks Dirichlet particles. */
void orc_cg(orc_system* sys, const orc_step_params* prm, orc_step_stats* st);
/* Mixed-precision solve (precision = 1): iterative refinement in FP64 with
 * the FP64 system's residual around a block-Jacobi PCG on FP32 vectors and
 * FP32 blocks whose dot products and norms are FP64 (orc_dot's tree over
 * float operands). */
void orc_cg_mixed(orc_system* sys, const orc_step_params* prm, orc_step_stats* st);
/* block-Jacobi inverse, then val32 / diag32 / minv32 = float(val / diag / minv) */
void orc_mixed_matrix(orc_system* sys);
/* val32 / diag32 / minv32 = float(val / diag / minv) */
void orc_round_blocks(orc_system* sys);
/* y = A32 x in FP64 arithmetic (FP64 diagonal, float bonds widened) */
void orc_spmv_wide(const orc_system* sys, const double* x4, double* y4);
double orc_dot_f32(int32_t n, const float* a4, const float* b4);
/* core/implicit_euler.cuh:90-176 / core/newmark_beta.cuh:89-199 */
void orc_update(orc_system* sys, const orc_step_params* prm);
/* core/simulator.cuh:602-768 collidePDParticles restated on a uniform
 * grid: fc[p] = sum over q != p with r < Rc of Kc (r-Rc)^2/V (p-q)/r, q of
 * another body (contact_mode 0) or any (contact_mode 1, the reference's
 * pair set) */
void orc_contact(orc_system* sys, const orc_step_params* prm);
/* core/simulator.cu:297-467: contact, assemble, s <- s_new, solve,
 * v <- sol, update.  Returns 0, or 1 if CG did not converge. */
int32_t orc_step(orc_system* sys, const orc_step_params* prm,
                 orc_step_stats* st);

#ifdef __cplusplus
}
#endif

#endif /* PD_ORACLE_H */
