// parameters.h -- the params.txt surface of the reference (core/parameters.h)
// with the same structs, field names and types; fields added for lattice
// scenes, SI time steps, solver control and contact sit at the end of each
// struct so the reference's fields keep their offsets.
#ifndef PD_HOST_PARAMETERS_H
#define PD_HOST_PARAMETERS_H

#include <stdint.h>

#include "definitions.h"

struct SimulationParameters {
  // read from file (core/parameters.h:16-87)
  int num_total_particle;
  int num_sph_particle;
  int num_pd_particle;

  int num_cell_x;
  int num_cell_y;
  int num_cell_z;
  int num_cells;
  int num_clists;

  real_t cell_size;

  real_t sph_kernel_coeff;
  real_t pd_kernel_coeff;
  real_t sph_particle_mass;
  real_t pd_particle_mass;
  real_t sph_sph_viscosity;
  real_t sph_boundary_viscosity;
  real_t sph_pd_slip;

  real_t boundary_min_x;
  real_t boundary_min_y;
  real_t boundary_min_z;
  real_t boundary_max_x;
  real_t boundary_max_y;
  real_t boundary_max_z;

  real_t pd_min_x;
  real_t pd_min_y;
  real_t pd_min_z;
  real_t pd_max_x;
  real_t pd_max_y;
  real_t pd_max_z;

  int boundaryPlaneSize;
  int boundaryPlaneBottomIndex;
  int boundaryPlaneBottomSize;
  int boundaryPlaneTopIndex;
  int boundaryPlaneTopSize;
  int boundaryPlaneFrontIndex;
  int boundaryPlaneFrontSize;
  int boundaryPlaneBackIndex;
  int boundaryPlaneBackSize;
  int boundaryPlaneLeftSideIndex;
  int boundaryPlaneLeftSideSize;
  int boundaryPlaneRightSideIndex;
  int boundaryPlaneRightSideSize;

  int pd_max_num_bonds;
  int pd_time_step;

  real_t sph_kernel_smooth_length;
  real_t sph_kernel_smooth_length_squared;
  real_t sph_particle_radius;
  real_t sph_kernel_poly6;
  real_t sph_kernel_spiky;
  real_t sph_rest_density;

  real_t pd_particle_radius;
  real_t pd_horizon;
  real_t pd_C_times_V;
  real_t sph_pd_min_distance;

  real_t clockScale;
  real_t scaleFactor;

  // ---- added ----
  real_t pd_dt;             // > 0: time step in seconds (else pd_time_step * 2^-23)
  real_t pd_bond_stiffness; // > 0: bond stiffness k [N/m], c_bond = k * spacing
  real_t pd_gravity;        // 9.8, along -y
  int pd_phase_in;          // PHASEIN; 0 disables the phase-in
  int num_pd_bodies;        // 1, or 2 with a projectile
  real_t pd_c_bond;         // force per unit stretch of body 0 (C*K or k*h)
  real_t projectile_spacing;
  real_t projectile_particle_radius;
  real_t projectile_horizon;
  real_t projectile_c_bond;
};

struct RunningParameters {
  RunningParameters();

  int adaptive_integration;
  int gpu_device;
  int step_per_frame;
  int step_per_state;
  int start_step;
  int final_step;

  int wall_thickness;
  int tube_radius;
  float emiter_x;
  float emiter_y;
  float sph_initial_velocity;
  float mesh_translation_x;
  float mesh_translation_y;
  float mesh_translation_z;
  float wall_translation_z;

  float pd_stretch_limit_s0;
  float pd_initial_velocity;

  int dump_activity;
  int dump_density;
  int dump_velocity;
  int dump_pd_bond_list_top;

  char obj_file[512];
  char saving_path[512];

  Integrator integrator;

  // ---- added ----
  int cg_preconditioner;        // 0 identity (reference), 1 block Jacobi, 2 learned
  char learned_network[512];    // parameter file of the learned preconditioner
  double cg_relative_tolerance; // CG_RELATIVE_TOLERANCE
  int cg_max_iterations;        // MAX_CG_ITERATION
  int cg_precision;             // 0 FP64, 1 float inner solve + FP64 refinement
  double cg_inner_rtol;         // precision 1: inner reduction (0: 1e-5)
  int cg_max_refine;            // precision 1: refinement limit (0: 20)
  int scene_lattice;            // 1: box lattice scene, 2: batch of lattice specimens
                                //    (0: the reference's mesh fill)
  int lattice_nx, lattice_ny, lattice_nz;
  double lattice_spacing;       // [m]
  double lattice_jitter;        // uniform jitter, fraction of the spacing
  unsigned lattice_seed;
  double lattice_origin_x, lattice_origin_y, lattice_origin_z;
  int bc_fix_zmin_layers;       // bottom z-layers held at v = 0
  int bc_move_zmax_layers;      // top z-layers moved at bc_move_zmax_velocity
  double bc_move_zmax_velocity; // [m/s] along +z
  int projectile_particles;     // 0: no projectile
  double projectile_radius;     // [m]
  double projectile_speed;      // [m/s] toward the top face (-z)
  double projectile_gap;        // [m] between sphere and top face
  unsigned projectile_seed;
  double contact_radius;        // Rc [m]; 0 disables contact
  double contact_stiffness;     // Kc
  double contact_volume;        // divisor of Kc (PD_V in the reference)
  int contact_mode;             // 0 between bodies, 1 all pairs (contact_pairs all)
  int walls;                    // 1: clamp to the boundary box (-1: mesh 1, lattice 0)
  unsigned stretch_seed;        // srand seed of getstretch (1973)
  int print_steps;              // 1: one JSON line per step on stdout
  // scene batch: independent tension specimens
  int batch_count;
  unsigned batch_seed;
  int batch_min_particles, batch_max_particles;  // log-uniform
  double batch_max_aspect;                       // long axis z : x = y, uniform in [1, max]
  double batch_s0_min, batch_s0_max;             // s0 per specimen, uniform
  int cg_matrix_free;           // cg_operator: 1 the FP64 SpMV recomputes the bond blocks,
                                // 2 (auto) the cheaper of stored / recomputed per step
};

class SystemParameter {
 public:
  explicit SystemParameter(const char* _paramFile);

  SimulationParameters& getSimulationParameters();
  RunningParameters& getRunningParameters();
  bool ok() const { return ok_; }

 private:
  void initParameters();
  void loadParametersFromFile(const char* _paramFile);
  void calculateDependentParameters();

  SimulationParameters simParams;
  RunningParameters runningParams;
  bool ok_ = false;
};

#endif
