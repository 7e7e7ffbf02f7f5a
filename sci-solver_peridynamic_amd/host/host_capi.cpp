// host_capi.cpp -- C entry points over the host classes for scripted
// callers (Python tests and the benchmark): parse a params file, build its
// scene, describe the derived run settings.  No device code.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "exact_predicates.h"
#include "mesh_fill.h"
#include "parameters.h"
#include "scene.h"

extern "C" {

// Parses params_path and builds its scene.  With null arrays only the
// particle count is returned in *n_out; otherwise the arrays (sized for
// *n_out particles) receive positions, velocities, body, Dirichlet flag and
// prescribed velocity.  Returns 0 or -1 (message on stderr).
int pdh_scene(const char* params_path, int32_t* n_out, double* x4, double* v4, int32_t* body,
              int32_t* bc, double* vbc) {
  SystemParameter sp(params_path);
  if (!sp.ok()) return -1;
  SimulationParameters& s = sp.getSimulationParameters();
  RunningParameters& r = sp.getRunningParameters();
  Scene scene(s, r, Scene::REGULAR_GRID);
  if (!scene.ok()) {
    fprintf(stderr, "%s\n", scene.error());
    return -1;
  }
  const int n = s.num_pd_particle;
  if (n_out) *n_out = n;
  if (!x4) return 0;
  std::vector<int> act(n);
  scene.initPeridynamicsParticles(act.data(), (real4_t*)x4, (real4_t*)v4);
  scene.initPeridynamicsBoundary(body, bc, (real4_t*)vbc);
  return 0;
}

// Writes a JSON object with the derived run settings of params_path into
// buf: materials per body, time step, step parameters, s0, step counts.
// Returns the JSON length, or -1.
int pdh_describe(const char* params_path, char* buf, int32_t len) {
  SystemParameter sp(params_path);
  if (!sp.ok()) return -1;
  SimulationParameters& s = sp.getSimulationParameters();
  RunningParameters& r = sp.getRunningParameters();
  Scene scene(s, r, Scene::REGULAR_GRID);
  const bool scene_ok = scene.ok();
  const double dt = s.pd_dt > 0.0 ? s.pd_dt * ((double)s.pd_time_step / (double)TIMESTEP_PD)
                                  : s.pd_time_step * SYSTEM_TIMESTEP_BASE;
  char tmp[8192];
  int w = snprintf(
      tmp, sizeof(tmp),
      "{\"num_pd_particle\": %d, \"scene_ok\": %d, \"num_bodies\": %d, "
      "\"materials\": [[%.17g, %.17g, %.17g, %.17g], [%.17g, %.17g, %.17g, %.17g]], "
      "\"dt\": %.17g, \"gravity\": %.17g, \"phase_in\": %d, \"integrator\": %d, "
      "\"precond\": %d, \"rtol\": %.17g, \"max_iter\": %d, \"walls\": %d, "
      "\"bmin\": [%.17g, %.17g, %.17g], \"bmax\": [%.17g, %.17g, %.17g], "
      "\"wall_radius\": %.17g, \"restitution\": %.17g, \"contact_radius\": %.17g, "
      "\"contact_stiffness\": %.17g, \"contact_volume\": %.17g, \"s0\": %.9g, "
      "\"stretch_seed\": %u, \"final_step\": %d, \"adaptive_integration\": %d, "
      "\"scale_factor\": %.17g, \"projectile_spacing\": %.17g, \"precision\": %d, "
      "\"inner_rtol\": %.17g, \"max_refine\": %d, \"contact_mode\": %d, "
      "\"matrix_free\": %d, \"cell_size\": %.17g, \"c_times_v\": %.17g}",
      s.num_pd_particle, scene_ok ? 1 : 0, s.num_pd_bodies, s.pd_horizon, s.pd_particle_radius,
      s.pd_c_bond, s.pd_particle_mass, s.projectile_horizon, s.projectile_particle_radius,
      s.projectile_c_bond, s.pd_particle_mass, dt, s.pd_gravity, s.pd_phase_in,
      r.integrator == IMPLICIT_EULER ? 0 : 1, r.cg_preconditioner, r.cg_relative_tolerance,
      r.cg_max_iterations, r.walls, s.boundary_min_x, s.boundary_min_y, s.boundary_min_z,
      s.boundary_max_x, s.boundary_max_y, s.boundary_max_z, s.pd_particle_radius, RESTITUTION,
      r.contact_radius, r.contact_stiffness, r.contact_volume, (double)r.pd_stretch_limit_s0,
      r.stretch_seed, r.final_step, r.adaptive_integration, s.scaleFactor, s.projectile_spacing,
      r.cg_precision, r.cg_inner_rtol > 0.0 ? r.cg_inner_rtol : 1e-5,
      r.cg_max_refine > 0 ? r.cg_max_refine : 20, r.contact_mode,
      r.cg_precision == 0 ? r.cg_matrix_free : 0, s.cell_size, s.pd_C_times_V);
  if (w < 0 || w >= len || !buf) return -1;
  memcpy(buf, tmp, (size_t)w + 1);
  return w;
}

// getstretch() of the host path (C library rand(), core/simulator.cu:802-864)
void pdh_stretch_init(int32_t n, float s0, uint32_t seed, double* out) {
  srand(seed);
  for (int32_t i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int k = 0; k < 12; ++k) sum += rand() / (double)RAND_MAX;
    sum -= 6.0;
    out[i] = ((sum * 0.01 + 1.0) * (double)s0);
  }
}

// The OBJ reader of mesh-filled scenes.  With null arrays only the counts
// are returned.  Returns 0 or -1 (message on stderr).
int pdh_load_obj(const char* path, int32_t* nv, int32_t* nf, float* v, int32_t* f) {
  ObjMesh m;
  std::string err;
  if (!load_obj(path, m, err)) {
    fprintf(stderr, "%s\n", err.c_str());
    return -1;
  }
  *nv = (int32_t)(m.v.size() / 3);
  *nf = (int32_t)(m.f.size() / 3);
  if (v) memcpy(v, m.v.data(), m.v.size() * sizeof(float));
  if (f) memcpy(f, m.f.data(), m.f.size() * sizeof(int32_t));
  return 0;
}

// Exact point-in-mesh test of n points against a triangle mesh.
void pdh_mesh_inside(int32_t nv, const double* xyz, int32_t nt, const int32_t* tri, int32_t n,
                     const double* pts, int32_t* out) {
  const std::vector<double> x(xyz, xyz + 3 * (size_t)nv);
  const std::vector<int> t(tri, tri + 3 * (size_t)nt);
  MeshInside m(x, t);
  for (int32_t i = 0; i < n; ++i) out[i] = m.inside(pts + 3 * (size_t)i) ? 1 : 0;
}

// Signs of orient2d(a, b, c) and orient3d(a, b, c, d) for n point tuples.
void pdh_orient(int32_t n, const double* a, const double* b, const double* c, const double* d,
                int32_t* o2, int32_t* o3) {
  for (int32_t i = 0; i < n; ++i) {
    o2[i] = exact::orient2d(a + 3 * i, b + 3 * i, c + 3 * i);
    o3[i] = exact::orient3d(a + 3 * i, b + 3 * i, c + 3 * i, d + 3 * i);
  }
}

// Activity, per-particle s0 and the specimen table of the params file's
// scene.  With null arrays only *n and *nspec are returned.  A scene without
// specimens reports nspec = 0 and s0 = pd_stretch_limit_s0 for every particle.
int pdh_scene_extra(const char* params_path, int32_t* n, int32_t* nspec, int32_t* act, float* s0,
                    int32_t* first, int32_t* count) {
  SystemParameter sp(params_path);
  if (!sp.ok()) return -1;
  SimulationParameters& s = sp.getSimulationParameters();
  RunningParameters& r = sp.getRunningParameters();
  Scene scene(s, r, Scene::REGULAR_GRID);
  if (!scene.ok()) {
    fprintf(stderr, "%s\n", scene.error());
    return -1;
  }
  *n = s.num_pd_particle;
  *nspec = scene.numSpecimens();
  if (!act) return 0;
  std::vector<real4_t> x(*n), v(*n);
  scene.initPeridynamicsParticles(act, x.data(), v.data());
  const std::vector<float>& s0v = scene.particleS0();
  for (int32_t i = 0; i < *n; ++i) s0[i] = s0v.empty() ? r.pd_stretch_limit_s0 : s0v[i];
  for (int32_t k = 0; k < *nspec; ++k) {
    first[k] = scene.specimenFirst()[k];
    count[k] = scene.specimenCount()[k];
  }
  return 0;
}

// getstretch() with a stretch limit per particle, one srand(seed) first.
void pdh_stretch_init_each(int32_t n, const float* s0, uint32_t seed, double* out) {
  srand(seed);
  for (int32_t i = 0; i < n; ++i) {
    double sum = 0.0;
    for (int k = 0; k < 12; ++k) sum += rand() / (double)RAND_MAX;
    sum -= 6.0;
    out[i] = ((sum * 0.01 + 1.0) * (double)s0[i]);
  }
}

}  // extern "C"
