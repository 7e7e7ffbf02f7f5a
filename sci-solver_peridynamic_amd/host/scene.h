// scene.h -- particle generation behind the reference's Scene interface
// (examples/scene.h:28-68).  The reference fills an OBJ mesh; this engine
// builds box-lattice specimens and an optional spherical projectile.
#ifndef PD_HOST_SCENE_H
#define PD_HOST_SCENE_H

#include <vector>

#include "definitions.h"
#include "parameters.h"

class Scene {
 public:
  enum ParticleArrangement { REGULAR_GRID, JITTERED_GRID };

  // examples/scene.cpp:18-136: computes the particle count into
  // simParams.num_pd_particle (and num_clists, num_total_particle) and the
  // projectile's material into simParams.projectile_*.
  Scene(SimulationParameters& simParams, RunningParameters& runningParams,
        ParticleArrangement arrangement = Scene::REGULAR_GRID);

  // examples/scene.cpp:360-395: positions (w = 0), activity and velocity
  // of all particles (specimen first, then the projectile).
  void initPeridynamicsParticles(int* pd_activity, real4_t* pd_position,
                                 real4_t* pd_velocity);
  // Added: body index, Dirichlet flag and prescribed velocity per particle.
  void initPeridynamicsBoundary(int* pd_body, int* pd_bc, real4_t* pd_bc_velocity);

  bool ok() const { return ok_; }
  const char* error() const { return err_; }

 private:
  void buildLattice();
  void buildProjectile();

  SimulationParameters& simParams_;
  RunningParameters& runningParams_;
  ParticleArrangement arrangement_;
  std::vector<real4_t> pos_;
  std::vector<real4_t> vel_;
  std::vector<int> body_;
  std::vector<int> bc_;
  std::vector<real4_t> vbc_;
  bool ok_ = false;
  const char* err_ = "";
};

// Counter-based uniform deviate in [0, 1): the splitmix64 finaliser of
// seed * 0x9E3779B97F4A7C15 + index, top 53 bits.  Order-independent, so a
// lattice point's jitter does not depend on how many points precede it.
double scene_uniform(unsigned seed, unsigned long long index);

#endif
