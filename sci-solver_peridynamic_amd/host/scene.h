// scene.h -- particle generation behind the reference's Scene interface
// (examples/scene.h:28-68): an OBJ mesh filled with a lattice (the
// reference's scene), or box-lattice specimens with an optional spherical
// projectile (scene lattice).
#ifndef PD_HOST_SCENE_H
#define PD_HOST_SCENE_H

#include <string>
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

  // Added (scene batch): independent specimens.  Specimen k owns rows
  // [first[k], first[k] + count[k]); the rows between specimens are
  // INVISIBLE padding (each specimen starts on a 256-row boundary).
  int numSpecimens() const { return (int)spec_first_.size(); }
  const std::vector<int>& specimenFirst() const { return spec_first_; }
  const std::vector<int>& specimenCount() const { return spec_count_; }
  // stretch limit s0 per particle (empty: pd_stretch_limit_s0 for all)
  const std::vector<float>& particleS0() const { return s0_; }

  bool ok() const { return ok_; }
  const char* error() const { return err_; }

 private:
  void buildLattice();
  void buildProjectile();
  bool buildMesh();
  bool buildBatch();

  static constexpr double kMeshJitter = 0.01;  // JITTER, examples/scene.h:20

  SimulationParameters& simParams_;
  RunningParameters& runningParams_;
  ParticleArrangement arrangement_;
  std::vector<real4_t> pos_;
  std::vector<real4_t> vel_;
  std::vector<int> body_;
  std::vector<int> bc_;
  std::vector<real4_t> vbc_;
  std::vector<int> act_;          // empty: all ACTIVE
  std::vector<int> spec_first_, spec_count_;
  std::vector<float> s0_;
  std::vector<double> mesh_xyz_;  // scaled mesh vertices (double)
  std::vector<int> mesh_tri_;
  bool ok_ = false;
  const char* err_ = "";
  std::string err_msg_;
};

// Counter-based uniform deviate in [0, 1): the splitmix64 finaliser of
// seed * 0x9E3779B97F4A7C15 + index, top 53 bits.  Order-independent, so a
// lattice point's jitter does not depend on how many points precede it.
double scene_uniform(unsigned seed, unsigned long long index);

#endif
