// parameters.cpp -- params.txt parser with the reference's semantics
// (core/parameters.cu:89-347): one "key value" pair per line, trailing
// whitespace trimmed, empty lines and lines containing "//" skipped,
// unknown keys ignored; then the dependent values (core/parameters.cu:
// 351-453).  Keys added by this engine are listed in the DESIGN notes of
// the package and parsed the same way.
#include "parameters.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>

RunningParameters::RunningParameters()
    : adaptive_integration(0),
      gpu_device(0),
      step_per_frame(1),
      step_per_state(0),
      start_step(0),
      final_step(1),
      wall_thickness(20),
      tube_radius(0),
      emiter_x(0.f),
      emiter_y(0.f),
      sph_initial_velocity(1.0f),
      mesh_translation_x(0.0f),
      mesh_translation_y(0.0f),
      mesh_translation_z(0.0f),
      wall_translation_z(0.0f),
      pd_stretch_limit_s0(0.0f),
      // never initialised in the reference (core/parameters.h:92-104); a
      // non-zero value puts a constant into the w lane of the right-hand
      // side, which the solver cannot match, so it is 0 here
      pd_initial_velocity(0.0f),
      dump_activity(0),
      dump_density(0),
      dump_velocity(0),
      dump_pd_bond_list_top(0),
      integrator(IMPLICIT_EULER),
      cg_preconditioner(0),
      cg_relative_tolerance(CG_RELATIVE_TOLERANCE),
      cg_max_iterations(MAX_CG_ITERATION),
      cg_precision(0),
      cg_inner_rtol(0.0),
      cg_max_refine(0),
      scene_lattice(0),
      lattice_nx(0),
      lattice_ny(0),
      lattice_nz(0),
      lattice_spacing(0.0),
      lattice_jitter(0.0),
      lattice_seed(1973u),
      lattice_origin_x(0.0),
      lattice_origin_y(0.0),
      lattice_origin_z(0.0),
      bc_fix_zmin_layers(0),
      bc_move_zmax_layers(0),
      bc_move_zmax_velocity(0.0),
      projectile_particles(0),
      projectile_radius(0.0),
      projectile_speed(0.0),
      projectile_gap(-1.0),
      projectile_seed(2024u),
      contact_radius(0.0),
      contact_stiffness(0.0),
      contact_volume(1.0),
      contact_mode(0),
      walls(-1),
      stretch_seed(1973u),
      print_steps(0),
      batch_count(0),
      batch_seed(2025u),
      batch_min_particles(5000),
      batch_max_particles(60000),
      batch_max_aspect(4.0),
      batch_s0_min(0.01),
      batch_s0_max(0.03),
      cg_matrix_free(0) {
  obj_file[0] = 0;
  learned_network[0] = 0;
  saving_path[0] = 0;
}

SystemParameter::SystemParameter(const char* _paramFile) {
  initParameters();
  loadParametersFromFile(_paramFile);
}

SimulationParameters& SystemParameter::getSimulationParameters() { return simParams; }
RunningParameters& SystemParameter::getRunningParameters() { return runningParams; }

// core/parameters.cu:40-86
void SystemParameter::initParameters() {
  memset(&simParams, 0, sizeof(simParams));
  simParams.num_cell_x = 1;
  simParams.num_cell_y = 1;
  simParams.num_cell_z = 1;
  simParams.num_cells = 1;
  simParams.sph_kernel_coeff = 4.001f;
  simParams.pd_kernel_coeff = 6.001;
  simParams.sph_particle_mass = 1.0f;
  simParams.pd_particle_mass = 1.0f;
  simParams.sph_sph_viscosity = 1e-5;
  simParams.sph_pd_slip = 1.0;
  simParams.sph_boundary_viscosity = 1e-5;
  simParams.boundary_min_x = 0.0f;
  simParams.boundary_min_y = 0.0f;
  simParams.boundary_min_z = 0.0f;
  simParams.boundary_max_x = 1.0f;
  simParams.boundary_max_y = 1.0f;
  simParams.boundary_max_z = 1.0f;
  simParams.clockScale = 0.0f;
  simParams.pd_gravity = 9.8;
  simParams.pd_phase_in = PHASEIN;
  simParams.num_pd_bodies = 1;
}

namespace {

bool parse_bool_int(const std::string& v) { return atoi(v.c_str()) != 0; }

}  // namespace

void SystemParameter::loadParametersFromFile(const char* _paramFile) {
  std::ifstream inFile(_paramFile);
  if (!inFile.is_open()) {
    fprintf(stderr, "Could not open parameter file: %s\n", _paramFile);
    ok_ = false;
    return;
  }
  std::string line, name, value;
  SimulationParameters& s = simParams;
  RunningParameters& r = runningParams;
  while (std::getline(inFile, line)) {
    line.erase(line.find_last_not_of(" \n\r\t") + 1);
    if (line.empty()) continue;
    if (line.find("//") != std::string::npos) continue;
    std::istringstream iss(line);
    name.clear();
    value.clear();
    iss >> name >> value;
    const char* v = value.c_str();
    // reference keys (core/parameters.cu:121-340)
    if (name == "num_sph_particle") s.num_sph_particle = atoi(v);
    if (name == "num_pd_particle") s.num_pd_particle = atoi(v);
    if (name == "radius_pd_over_sph") s.pd_particle_radius = atof(v);
    if (name == "boundary_min_x") s.boundary_min_x = atof(v);
    if (name == "boundary_min_y") s.boundary_min_y = atof(v);
    if (name == "boundary_min_z") s.boundary_min_z = atof(v);
    if (name == "boundary_max_x") s.boundary_max_x = atof(v);
    if (name == "boundary_max_y") s.boundary_max_y = atof(v);
    if (name == "boundary_max_z") s.boundary_max_z = atof(v);
    if (name == "sph_kernel_coeff") s.sph_kernel_coeff = atof(v);
    if (name == "pd_kernel_coeff") s.pd_kernel_coeff = atof(v);
    if (name == "sph_particle_mass") s.sph_particle_mass = atof(v);
    if (name == "pd_particle_mass") s.pd_particle_mass = atof(v);
    if (name == "sph_sph_viscosity") s.sph_sph_viscosity = atof(v);
    if (name == "sph_boundary_viscosity") s.sph_boundary_viscosity = atof(v);
    if (name == "sph_pd_slip") s.sph_pd_slip = atof(v);
    if (name == "adaptive_integration") r.adaptive_integration = atoi(v);
    if (name == "gpu_device") r.gpu_device = atoi(v);
    if (name == "steps_per_frame") r.step_per_frame = atoi(v);
    if (name == "tube_radius") r.tube_radius = atoi(v);
    if (name == "wall_thickness") r.wall_thickness = atoi(v);
    if (name == "mesh_translation_x") r.mesh_translation_x = (float)atof(v);
    if (name == "mesh_translation_y") r.mesh_translation_y = (float)atof(v);
    if (name == "sph_initial_velocity") r.sph_initial_velocity = (float)atof(v);
    if (name == "mesh_translation_z") r.mesh_translation_z = (float)atof(v);
    if (name == "pd_stretch_limit_s0") r.pd_stretch_limit_s0 = (float)atof(v);
    if (name == "pd_initial_velocity") r.pd_initial_velocity = (float)atof(v);
    if (name == "steps_per_state") r.step_per_state = atoi(v);
    if (name == "start_step") r.start_step = atoi(v);
    if (name == "final_step") r.final_step = atoi(v);
    if (name == "dump_density") r.dump_density = atoi(v);
    if (name == "dump_velocity") r.dump_velocity = atoi(v);
    if (name == "dump_activity") r.dump_activity = atoi(v);
    if (name == "dump_pd_bond_list_top") r.dump_pd_bond_list_top = atoi(v);
    if (name == "obj_file") snprintf(r.obj_file, sizeof(r.obj_file), "%s", v);
    if (name == "saving_path") snprintf(r.saving_path, sizeof(r.saving_path), "%s", v);
    if (name == "integrator") r.integrator = (value == "implicit_euler") ? IMPLICIT_EULER : NEWMARK_BETA;
    // keys added by this engine
    if (name == "pd_dt") s.pd_dt = atof(v);
    if (name == "pd_bond_stiffness") s.pd_bond_stiffness = atof(v);
    if (name == "pd_gravity") s.pd_gravity = atof(v);
    if (name == "pd_phase_in") s.pd_phase_in = atoi(v);
    if (name == "cg_preconditioner")
      r.cg_preconditioner = (value == "block_jacobi" || value == "1")
                                ? 1
                                : ((value == "learned" || value == "2") ? 2 : 0);
    if (name == "learned_network")
      snprintf(r.learned_network, sizeof(r.learned_network), "%s", v);
    if (name == "cg_relative_tolerance") r.cg_relative_tolerance = atof(v);
    if (name == "cg_max_iterations") r.cg_max_iterations = atoi(v);
    if (name == "cg_precision") r.cg_precision = (value == "mixed" || value == "1") ? 1 : 0;
    if (name == "cg_inner_rtol") r.cg_inner_rtol = atof(v);
    if (name == "cg_operator")
      r.cg_matrix_free = (value == "matrix_free" || value == "1") ? 1
                         : (value == "auto" || value == "2")      ? 2
                                                                  : 0;
    if (name == "cg_max_refine") r.cg_max_refine = atoi(v);
    if (name == "scene") r.scene_lattice = (value == "lattice") ? 1 : (value == "batch" ? 2 : 0);
    if (name == "batch_count") r.batch_count = atoi(v);
    if (name == "batch_seed") r.batch_seed = (unsigned)strtoul(v, nullptr, 10);
    if (name == "batch_min_particles") r.batch_min_particles = atoi(v);
    if (name == "batch_max_particles") r.batch_max_particles = atoi(v);
    if (name == "batch_max_aspect") r.batch_max_aspect = atof(v);
    if (name == "batch_s0_min") r.batch_s0_min = atof(v);
    if (name == "batch_s0_max") r.batch_s0_max = atof(v);
    if (name == "lattice_nx") r.lattice_nx = atoi(v);
    if (name == "lattice_ny") r.lattice_ny = atoi(v);
    if (name == "lattice_nz") r.lattice_nz = atoi(v);
    if (name == "lattice_spacing") r.lattice_spacing = atof(v);
    if (name == "lattice_jitter") r.lattice_jitter = atof(v);
    if (name == "lattice_seed") r.lattice_seed = (unsigned)strtoul(v, nullptr, 10);
    if (name == "lattice_origin_x") r.lattice_origin_x = atof(v);
    if (name == "lattice_origin_y") r.lattice_origin_y = atof(v);
    if (name == "lattice_origin_z") r.lattice_origin_z = atof(v);
    if (name == "bc_fix_zmin_layers") r.bc_fix_zmin_layers = atoi(v);
    if (name == "bc_move_zmax_layers") r.bc_move_zmax_layers = atoi(v);
    if (name == "bc_move_zmax_velocity") r.bc_move_zmax_velocity = atof(v);
    if (name == "projectile_particles") r.projectile_particles = atoi(v);
    if (name == "projectile_radius") r.projectile_radius = atof(v);
    if (name == "projectile_speed") r.projectile_speed = atof(v);
    if (name == "projectile_gap") r.projectile_gap = atof(v);
    if (name == "projectile_seed") r.projectile_seed = (unsigned)strtoul(v, nullptr, 10);
    if (name == "contact_radius") r.contact_radius = atof(v);
    if (name == "contact_stiffness") r.contact_stiffness = atof(v);
    if (name == "contact_volume") r.contact_volume = atof(v);
    if (name == "contact_pairs") r.contact_mode = (value == "all" || value == "1") ? 1 : 0;
    if (name == "walls") r.walls = parse_bool_int(value) ? 1 : 0;
    if (name == "stretch_seed") r.stretch_seed = (unsigned)strtoul(v, nullptr, 10);
    if (name == "print_steps") r.print_steps = atoi(v);
  }
  calculateDependentParameters();
  ok_ = true;
}

// core/parameters.cu:351-453 (mesh scenes, scaled units); lattice scenes
// keep SI units: scaleFactor 1, radius h/2, horizon = radius * pd_kernel_coeff,
// c_bond = k h.
void SystemParameter::calculateDependentParameters() {
  SimulationParameters& s = simParams;
  RunningParameters& r = runningParams;
  s.num_total_particle = s.num_sph_particle + s.num_pd_particle;
  s.num_cell_x = NUM_CELL_X;
  s.num_cell_y = NUM_CELL_Y;
  s.num_cell_z = NUM_CELL_Z;
  s.num_cells = NUM_CELLS;
  s.pd_time_step = TIMESTEP_PD;
  if (r.walls < 0) r.walls = r.scene_lattice ? 0 : 1;  // walls on for mesh scenes
  if (r.scene_lattice) {
    const double h = r.lattice_spacing;
    s.cell_size = h * s.pd_kernel_coeff * 0.5;
    s.scaleFactor = 1.0;
    s.pd_particle_radius = 0.5 * h;
    s.pd_horizon = s.pd_particle_radius * s.pd_kernel_coeff;
    s.pd_C_times_V = 0.0;
    s.pd_c_bond = s.pd_bond_stiffness * h;
    s.num_pd_bodies = r.projectile_particles > 0 ? 2 : 1;
    return;
  }
  // core/parameters.cu:360-363: fmaxf on the three extents (float rounding)
  const float ex = (float)((s.boundary_max_x - s.boundary_min_x) / NUM_CELL_X);
  const float ey = (float)((s.boundary_max_y - s.boundary_min_y) / NUM_CELL_Y);
  const float ez = (float)((s.boundary_max_z - s.boundary_min_z) / NUM_CELL_Z);
  const float mxy = ex > ey ? ex : ey;
  s.cell_size = mxy > ez ? mxy : ez;
  s.sph_kernel_smooth_length = s.cell_size;
  s.sph_particle_radius = s.cell_size / s.sph_kernel_coeff;
  s.pd_particle_radius *= s.sph_particle_radius;
  s.pd_horizon = s.pd_particle_radius * s.pd_kernel_coeff;
  s.pd_C_times_V = (18.0 / PD_PI / 2.0 / 1296.0 / s.pd_particle_radius);
  s.sph_kernel_poly6 = 315.0 / 64.0 / PD_PI;
  s.sph_kernel_spiky = 45.0 / PD_PI;
  s.scaleFactor = 2 * s.pd_particle_radius / PERIDYNAMICS_MIN_DISTANCE;
  s.boundary_min_x /= s.scaleFactor;
  s.boundary_max_x /= s.scaleFactor;
  s.boundary_min_y /= s.scaleFactor;
  s.boundary_max_y /= s.scaleFactor;
  s.boundary_min_z /= s.scaleFactor;
  s.boundary_max_z /= s.scaleFactor;
  s.sph_particle_radius /= s.scaleFactor;
  s.pd_particle_radius /= s.scaleFactor;
  s.sph_pd_min_distance = s.sph_particle_radius + s.pd_particle_radius;
  s.cell_size /= s.scaleFactor;
  s.pd_horizon /= s.scaleFactor;
  s.sph_kernel_smooth_length /= s.scaleFactor;
  s.sph_kernel_smooth_length_squared = s.sph_kernel_smooth_length * s.sph_kernel_smooth_length;
  s.boundaryPlaneSize = (int)ceil(3 * s.cell_size / (s.sph_particle_radius * 2));
  s.pd_c_bond = s.pd_C_times_V * PD_K;
  if (s.pd_bond_stiffness > 0.0) s.pd_c_bond = s.pd_bond_stiffness * 2.0 * s.pd_particle_radius;
  s.num_pd_bodies = 1;
}
