// simulator.h -- the reference's Simulator interface (core/simulator.h:32-57)
// on top of the C ABI of include/pd_solver.h.  Host mirrors hold the
// particle arrays the caller fills before makeReady() and reads after
// dataIO() downloads.
#ifndef PD_HOST_SIMULATOR_H
#define PD_HOST_SIMULATOR_H

#include <string>
#include <vector>

#include "data_io.h"
#include "definitions.h"
#include "parameters.h"
#include "pd_solver.h"

class Simulator {
 public:
  Simulator(int deviceID, SimulationParameters& simParams, RunningParameters& runningParams);
  ~Simulator();

  void makeReady();
  void advance();
  void printSimulationTime();

  bool finished();
  bool CGConvergence();

  int get_current_step() const;
  // added: continue a restored state from step `step` (DataIO::loadState)
  void resumeAt(int step) { current_step_ = step; }
  int* get_pd_activity();
  real4_t* get_pd_position();
  real4_t* get_pd_velocity();

  // added: independent specimens and per-particle stretch limits (scene
  // batch); call before makeReady()
  void setSpecimens(const std::vector<int>& first, const std::vector<int>& count) {
    spec_first_ = first;
    spec_count_ = count;
  }
  void setStretchLimits(const std::vector<float>& s0) { s0_ = s0; }

  // added: body index, Dirichlet flag / velocity (filled by the Scene)
  int* get_pd_body();
  int* get_pd_bc();
  real4_t* get_pd_bc_velocity();

  DataIO& dataIO();

  bool ok() const { return err_.empty(); }
  const std::string& error() const { return err_; }
  const pd_step_stats& lastStepStats() const { return last_; }
  const pd_bond_stats& bondStats() const { return bond_stats_; }
  pd_ctx* context() { return ctx_; }
  // time step of the next advance() in the engine's units
  double timestep() const;
  pd_step_params stepParams() const;

 protected:
  bool integratePeridynamics();
  void initPeridynamicsBonds();
  bool check(int rc, const char* what);

  int current_step_ = 0;
  int current_substep_ = 0;
  bool CG_convergent_ = true;

  SimulationParameters& hostSimParams_;
  RunningParameters& runningParams_;
  pd_ctx* ctx_ = nullptr;
  std::vector<int> activity_, body_, bc_;
  std::vector<real4_t> position_, velocity_, bc_velocity_;
  std::vector<double> stretch_;
  std::vector<int> spec_first_, spec_count_;
  std::vector<float> s0_;
  pd_step_stats last_ = {};
  pd_bond_stats bond_stats_ = {};
  double time_total_ms_ = 0.0, time_assemble_ms_ = 0.0, time_solve_ms_ = 0.0,
         time_contact_ms_ = 0.0;
  int steps_timed_ = 0;
  std::string err_;
  DataIO dataIO_;  // declared after the mirrors it reads
};

#endif
