// simulator.cpp -- Simulator (core/simulator.cu) above the C ABI.
#include "simulator.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>

Simulator::Simulator(int deviceID, SimulationParameters& simParams,
                     RunningParameters& runningParams)
    : hostSimParams_(simParams), runningParams_(runningParams), dataIO_(simParams, runningParams, *this) {
  const int n = std::max(0, hostSimParams_.num_pd_particle);
  activity_.assign(n, ACTIVE);
  body_.assign(n, 0);
  bc_.assign(n, 0);
  position_.assign(n, make_real4(0, 0, 0, 0));
  velocity_.assign(n, make_real4(0, 0, 0, 0));
  bc_velocity_.assign(n, make_real4(0, 0, 0, 0));
  stretch_.assign(n, 0.0);
  pd_config cfg = {};
  cfg.num_particles = n;
  cfg.device = deviceID;
  cfg.max_bonds = MAX_PD_BOND_COUNT;
  cfg.num_bodies = std::max(1, hostSimParams_.num_pd_bodies);
  const SimulationParameters& s = hostSimParams_;
  cfg.material[0] = {s.pd_horizon, s.pd_particle_radius, s.pd_c_bond, s.pd_particle_mass};
  if (cfg.num_bodies > 1)
    cfg.material[1] = {s.projectile_horizon, s.projectile_particle_radius, s.projectile_c_bond,
                       s.pd_particle_mass};
  if (n > 0) check(pd_ctx_create(&cfg, &ctx_), "pd_ctx_create");
}

Simulator::~Simulator() {
  if (ctx_) pd_ctx_destroy(ctx_);
}

bool Simulator::check(int rc, const char* what) {
  if (rc == PD_OK) return true;
  err_ = std::string(what) + ": " + pd_last_error(ctx_);
  fprintf(stderr, "%s\n", err_.c_str());
  return false;
}

double Simulator::timestep() const {
  // reference: pd_time_step * 2^-23 (implicit_euler.cuh:82); SI runs scale
  // pd_dt by the same adaptive factor pd_time_step / TIMESTEP_PD
  if (hostSimParams_.pd_dt > 0.0)
    return hostSimParams_.pd_dt * ((double)hostSimParams_.pd_time_step / (double)TIMESTEP_PD);
  return hostSimParams_.pd_time_step * SYSTEM_TIMESTEP_BASE;
}

pd_step_params Simulator::stepParams() const {
  const SimulationParameters& s = hostSimParams_;
  const RunningParameters& r = runningParams_;
  pd_step_params p = {};
  p.dt = timestep();
  p.clock_scale = s.clockScale;
  p.gravity = s.pd_gravity;
  p.integrator = r.integrator == IMPLICIT_EULER ? 0 : 1;
  p.precond = r.cg_preconditioner;
  p.rtol = r.cg_relative_tolerance;
  p.max_iter = r.cg_max_iterations;
  p.walls = r.walls;
  p.bmin[0] = s.boundary_min_x;
  p.bmin[1] = s.boundary_min_y;
  p.bmin[2] = s.boundary_min_z;
  p.bmax[0] = s.boundary_max_x;
  p.bmax[1] = s.boundary_max_y;
  p.bmax[2] = s.boundary_max_z;
  p.wall_radius = s.pd_particle_radius;
  p.restitution = RESTITUTION;
  p.contact_radius = r.contact_radius;
  p.contact_stiffness = r.contact_stiffness;
  p.contact_volume = r.contact_volume;
  p.contact_cell = r.contact_radius;
  p.contact_mode = r.contact_mode;
  p.track_gap = 0;
  p.timing = 1;
  p.precision = r.cg_precision;
  p.matrix_free = r.cg_precision == 0 ? r.cg_matrix_free : 0;
  p.inner_rtol = r.cg_inner_rtol;
  p.max_refine = r.cg_max_refine;
  p.inner_max_iter = 0;
  return p;
}

// core/simulator.cu:56-86: bonds and stretch limits, upload, zero solution
void Simulator::makeReady() {
  if (!ctx_ || !ok()) return;
  initPeridynamicsBonds();
}

// core/simulator.cu:825-943: x0 = x, s00 = s = getstretch(s0) after
// srand(1973) in particle order (the C library's rand(), as the reference
// calls it), then the horizon bond lists (built on the GPU).
void Simulator::initPeridynamicsBonds() {
  const int n = hostSimParams_.num_pd_particle;
  srand(runningParams_.stretch_seed);
  const bool each = s0_.size() == (size_t)n;
  for (int i = 0; i < n; ++i) {
    const double s0 = (double)(each ? s0_[i] : runningParams_.pd_stretch_limit_s0);
    double sum = 0.0;
    for (int k = 0; k < 12; ++k) sum += rand() / (double)RAND_MAX;
    sum -= 6.0;
    stretch_[i] = ((sum * 0.01 + 1.0) * s0);
  }
  std::vector<real4_t> zero(n, make_real4(0, 0, 0, 0));
  const size_t v4 = sizeof(real4_t) * (size_t)n, v1 = sizeof(double) * (size_t)n,
               i1 = sizeof(int) * (size_t)n;
  if (!check(pd_upload(ctx_, PD_FIELD_POSITION, position_.data(), v4), "upload position")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_ORIGINAL_POSITION, position_.data(), v4), "upload x0")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_VELOCITY, velocity_.data(), v4), "upload velocity")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_STRETCH, stretch_.data(), v1), "upload stretch")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_ORIGINAL_STRETCH, stretch_.data(), v1), "upload s00")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_ACTIVITY, activity_.data(), i1), "upload activity")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_BODY, body_.data(), i1), "upload body")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_BC_FLAG, bc_.data(), i1), "upload bc")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_BC_VELOCITY, bc_velocity_.data(), v4), "upload vbc")) return;
  if (!check(pd_upload(ctx_, PD_FIELD_SOLUTION, zero.data(), v4), "upload solution")) return;
  if (!check(pd_build_bonds(ctx_, &bond_stats_), "pd_build_bonds")) return;
  if (runningParams_.cg_preconditioner == 2) {
    // the learned preconditioner's parameters: one float per line
    std::vector<float> net;
    if (FILE* fp = fopen(runningParams_.learned_network, "r")) {
      float v;
      while (fscanf(fp, "%f", &v) == 1) net.push_back(v);
      fclose(fp);
    }
    if (!check(pd_learned_set(ctx_, net.data(), (int32_t)net.size()), "learned_network")) return;
  }
  hostSimParams_.pd_max_num_bonds = bond_stats_.max_bonds;
  if (!spec_first_.empty() &&
      !check(pd_set_specimens(ctx_, (int32_t)spec_first_.size(), spec_first_.data(),
                              spec_count_.data()),
             "pd_set_specimens"))
    return;
  printf("    Number of Peridynamics particle bonds: %d -> %d\n", bond_stats_.min_bonds,
         bond_stats_.max_bonds);
  printf("    Sum of bonds: %lld\n", (long long)bond_stats_.total_bonds);
  check(pd_backup(ctx_), "pd_backup");
}

bool Simulator::finished() {
  return (current_step_ > runningParams_.final_step) || !CG_convergent_ || !ok();
}

bool Simulator::CGConvergence() { return CG_convergent_; }
int Simulator::get_current_step() const { return current_step_; }
int* Simulator::get_pd_activity() { return activity_.data(); }
real4_t* Simulator::get_pd_position() { return position_.data(); }
real4_t* Simulator::get_pd_velocity() { return velocity_.data(); }
int* Simulator::get_pd_body() { return body_.data(); }
int* Simulator::get_pd_bc() { return bc_.data(); }
real4_t* Simulator::get_pd_bc_velocity() { return bc_velocity_.data(); }
DataIO& Simulator::dataIO() { return dataIO_; }

// core/simulator.cu:1159-1186
void Simulator::advance() {
  if (!ctx_ || !ok()) return;
  ++current_step_;
  const int phase = hostSimParams_.pd_phase_in;
  if (phase <= 0) {
    hostSimParams_.clockScale = 1.0;
  } else if (current_step_ <= phase) {
    hostSimParams_.clockScale = (real_t)current_step_ / (real_t)phase;
  }
  if (hostSimParams_.num_pd_particle <= 0) return;
  while (!integratePeridynamics()) {
    if (!ok()) return;
  }
}

// One attempt of core/simulator.cu:297-467 (implicit Euler) or 470-652
// (Newmark-beta, whose adaptive rollback is always on).  Returns false when
// the step must be retried with the halved time step.
bool Simulator::integratePeridynamics() {
  const bool adaptive =
      runningParams_.adaptive_integration || runningParams_.integrator != IMPLICIT_EULER;
  pd_step_params prm = stepParams();
  int rc;
  if (!adaptive) {
    rc = pd_step(ctx_, &prm, &last_);
  } else {
    if (!check(pd_assemble(ctx_, &prm, &last_), "pd_assemble")) return true;
    if (last_.broken_bonds > 0) {
      printf("        Has broken bonds!\n");
      if (hostSimParams_.pd_time_step > TIMESTEP_PD_MIN) {
        hostSimParams_.pd_time_step /= 2;
        printf("        Step back. Reduce time step to %d\n", hostSimParams_.pd_time_step);
        check(pd_restore(ctx_), "pd_restore");
        return false;
      }
      current_substep_ = 1;
    } else {
      if (current_substep_ > 0) ++current_substep_;
      if (current_substep_ >= MAX_SUBSTEP) {
        hostSimParams_.pd_time_step *= 2;
        printf("        Try to increase time step twice to %d\n", hostSimParams_.pd_time_step);
        current_substep_ = (hostSimParams_.pd_time_step >= TIMESTEP_PD) ? 0 : 1;
        // the reference fills A and b after this point with the doubled step
        // (core/simulator.cu:381-392, 430-446); the fill is fused into the
        // assembly here, so assemble again: no bond breaks on this path and
        // the force kernel does not depend on the step
        prm = stepParams();
        if (!check(pd_assemble(ctx_, &prm, &last_), "pd_assemble")) return true;
      }
    }
    if (!check(pd_backup(ctx_), "pd_backup")) return true;
    if (!check(pd_commit_stretch(ctx_), "pd_commit_stretch")) return true;
    pd_step_stats sst = last_;
    rc = pd_solve(ctx_, &prm, &sst);
    last_.cg_iterations = sst.cg_iterations;
    last_.converged = sst.converged;
    last_.residual_norm = sst.residual_norm;
    last_.b_norm = sst.b_norm;
    last_.ms_spmv_total = sst.ms_spmv_total;
    last_.spmv_launches = sst.spmv_launches;
    if (rc == PD_OK || rc == PD_ERR_NOT_CONVERGED) {
      const int rc2 = pd_update(ctx_, &prm);
      if (rc2 != PD_OK) rc = rc2;
    }
  }
  if (rc == PD_ERR_NOT_CONVERGED) {
    // core/simulator.cu:780-790
    printf("        CG Solver reached iteration limit %d before converging (%g final residual)\n",
           prm.max_iter, last_.residual_norm);
    CG_convergent_ = false;
  } else if (!check(rc, "step")) {
    return true;
  }
  time_contact_ms_ += last_.ms_contact;
  time_assemble_ms_ += last_.ms_assemble;
  time_solve_ms_ += last_.ms_solve;
  time_total_ms_ += last_.ms_contact + last_.ms_assemble + last_.ms_solve + last_.ms_update;
  ++steps_timed_;
  if (runningParams_.print_steps) {
    printf("{\"step\": %d, \"cg_iterations\": %d, \"converged\": %d, \"residual\": %.17g, "
           "\"b_norm\": %.17g, \"broken_bonds\": %lld, \"live_bonds\": %lld, \"dt\": %.17g}\n",
           current_step_, last_.cg_iterations, last_.converged, last_.residual_norm,
           last_.b_norm, (long long)last_.broken_bonds, (long long)last_.live_bonds, prm.dt);
    fflush(stdout);
  }
  return true;
}

// core/monitor.cu printSimulationTime: per-phase GPU time since the last call
void Simulator::printSimulationTime() {
  if (steps_timed_ == 0) return;
  printf("    Steps %d: contact %.3f ms, assemble %.3f ms, solve %.3f ms, total %.3f ms "
         "(%.3f ms/step)\n",
         steps_timed_, time_contact_ms_, time_assemble_ms_, time_solve_ms_, time_total_ms_,
         time_total_ms_ / steps_timed_);
  time_total_ms_ = time_assemble_ms_ = time_solve_ms_ = time_contact_ms_ = 0.0;
  steps_timed_ = 0;
}
