// scene.cpp -- box-lattice specimens and a seeded spherical projectile.
//
// Specimen: nx x ny x nz simple-cubic lattice, x fastest (the reference's
// fill order, examples/scene.cpp:536-540), spacing h, each coordinate
// displaced by (2u - 1) * jitter * h with u = scene_uniform(seed, 3 p + axis).
// Dirichlet layers: z-layers k < bc_fix_zmin_layers are held at v = 0,
// k >= nz - bc_move_zmax_layers move at (0, 0, bc_move_zmax_velocity).
//
// Projectile (body 1): the n points of a cubic lattice of spacing s closest
// to the sphere centre (ties by lattice index), s the largest spacing whose
// lattice has at least n points strictly inside radius R (bisection), then
// jittered like the specimen with projectile_seed.  Centre: above the middle
// of the top face by gap + R; velocity (0, 0, -speed).  Its material uses
// spacing s: radius s/2, horizon = radius * pd_kernel_coeff, c_bond = k s.
#include "scene.h"

#include <math.h>
#include <stdint.h>

#include <algorithm>

double scene_uniform(unsigned seed, unsigned long long index) {
  uint64_t z = (uint64_t)seed * 0x9E3779B97F4A7C15ull + index;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z = z ^ (z >> 31);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

Scene::Scene(SimulationParameters& simParams, RunningParameters& runningParams,
             ParticleArrangement arrangement)
    : simParams_(simParams), runningParams_(runningParams), arrangement_(arrangement) {
  if (!runningParams_.scene_lattice) {
    err_ = "mesh-filled scenes are not available in this engine; use 'scene lattice'";
    return;
  }
  const RunningParameters& r = runningParams_;
  if (r.lattice_nx <= 0 || r.lattice_ny <= 0 || r.lattice_nz <= 0 || !(r.lattice_spacing > 0.0)) {
    err_ = "lattice_nx/ny/nz and lattice_spacing must be positive";
    return;
  }
  buildLattice();
  if (r.projectile_particles > 0) {
    if (!(r.projectile_radius > 0.0)) {
      err_ = "projectile_radius must be positive";
      return;
    }
    buildProjectile();
  }
  simParams_.num_pd_particle = (int)pos_.size();
  simParams_.num_total_particle = simParams_.num_sph_particle + simParams_.num_pd_particle;
  // examples/scene.cpp:122-127
  simParams_.num_clists = simParams_.num_pd_particle;
  while (simParams_.num_clists % 8 != 0) simParams_.num_clists++;
  ok_ = true;
}

void Scene::buildLattice() {
  const RunningParameters& r = runningParams_;
  const double h = r.lattice_spacing;
  const double jit = arrangement_ == JITTERED_GRID || r.lattice_jitter > 0.0 ? r.lattice_jitter : 0.0;
  const int nx = r.lattice_nx, ny = r.lattice_ny, nz = r.lattice_nz;
  const size_t n = (size_t)nx * ny * nz;
  pos_.resize(n);
  vel_.assign(n, make_real4(0, 0, 0, 0));
  body_.assign(n, 0);
  bc_.assign(n, 0);
  vbc_.assign(n, make_real4(0, 0, 0, 0));
  const double org[3] = {r.lattice_origin_x, r.lattice_origin_y, r.lattice_origin_z};
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t p = ((size_t)k * ny + j) * nx + i;
        const int idx[3] = {i, j, k};
        double c[3];
        for (int a = 0; a < 3; ++a) {
          const double u = scene_uniform(r.lattice_seed, 3ull * p + a);
          c[a] = org[a] + idx[a] * h + (2.0 * u - 1.0) * jit * h;
        }
        pos_[p] = make_real4(c[0], c[1], c[2], 0.0);
        if (k < r.bc_fix_zmin_layers) {
          bc_[p] = 1;
        } else if (k >= nz - r.bc_move_zmax_layers) {
          bc_[p] = 1;
          vbc_[p] = make_real4(0.0, 0.0, r.bc_move_zmax_velocity, 0.0);
          vel_[p] = vbc_[p];
        }
      }
}

namespace {

int64_t count_inside(double s, double R) {
  const int m = (int)floor(R / s) + 1;
  int64_t c = 0;
  for (int k = -m; k <= m; ++k)
    for (int j = -m; j <= m; ++j)
      for (int i = -m; i <= m; ++i) {
        const double d = sqrt((double)i * i * s * s + (double)j * j * s * s + (double)k * k * s * s);
        if (d < R) ++c;
      }
  return c;
}

}  // namespace

void Scene::buildProjectile() {
  const RunningParameters& r = runningParams_;
  const double R = r.projectile_radius;
  const int64_t want = r.projectile_particles;
  // largest spacing with at least `want` lattice points inside the sphere
  double lo = R * 1e-3, hi = 2.0 * R;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_inside(mid, R) >= want) lo = mid;
    else hi = mid;
  }
  const double s = lo;
  struct Cand {
    double d;
    int64_t key;
    int i, j, k;
  };
  std::vector<Cand> cand;
  const int m = (int)floor(R / s) + 1;
  for (int k = -m; k <= m; ++k)
    for (int j = -m; j <= m; ++j)
      for (int i = -m; i <= m; ++i) {
        const double d = sqrt((double)i * i * s * s + (double)j * j * s * s + (double)k * k * s * s);
        if (d < R) {
          const int64_t key = ((int64_t)(k + m) * (2 * m + 1) + (j + m)) * (2 * m + 1) + (i + m);
          cand.push_back({d, key, i, j, k});
        }
      }
  std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) {
    return a.d < b.d || (a.d == b.d && a.key < b.key);
  });
  cand.resize((size_t)want);
  std::sort(cand.begin(), cand.end(), [](const Cand& a, const Cand& b) { return a.key < b.key; });
  const double h = r.lattice_spacing;
  const double top = r.lattice_origin_z + (r.lattice_nz - 1) * h;
  const double gap = r.projectile_gap >= 0.0 ? r.projectile_gap
                                             : simParams_.pd_horizon + h;  // default: beyond a horizon
  const double cx = r.lattice_origin_x + 0.5 * (r.lattice_nx - 1) * h;
  const double cy = r.lattice_origin_y + 0.5 * (r.lattice_ny - 1) * h;
  const double cz = top + gap + R;
  const double jit = r.lattice_jitter;
  for (size_t t = 0; t < cand.size(); ++t) {
    const int idx[3] = {cand[t].i, cand[t].j, cand[t].k};
    const double org[3] = {cx, cy, cz};
    double c[3];
    for (int a = 0; a < 3; ++a) {
      const double u = scene_uniform(r.projectile_seed, 3ull * t + a);
      c[a] = org[a] + idx[a] * s + (2.0 * u - 1.0) * jit * s;
    }
    pos_.push_back(make_real4(c[0], c[1], c[2], 0.0));
    vel_.push_back(make_real4(0.0, 0.0, -r.projectile_speed, 0.0));
    body_.push_back(1);
    bc_.push_back(0);
    vbc_.push_back(make_real4(0, 0, 0, 0));
  }
  simParams_.projectile_spacing = s;
  simParams_.projectile_particle_radius = 0.5 * s;
  simParams_.projectile_horizon = 0.5 * s * simParams_.pd_kernel_coeff;
  simParams_.projectile_c_bond = simParams_.pd_bond_stiffness * s;
  simParams_.num_pd_bodies = 2;
}

void Scene::initPeridynamicsParticles(int* pd_activity, real4_t* pd_position,
                                      real4_t* pd_velocity) {
  for (size_t p = 0; p < pos_.size(); ++p) {
    if (pd_activity) pd_activity[p] = ACTIVE;
    if (pd_position) pd_position[p] = pos_[p];
    if (pd_velocity) pd_velocity[p] = vel_[p];
  }
}

void Scene::initPeridynamicsBoundary(int* pd_body, int* pd_bc, real4_t* pd_bc_velocity) {
  for (size_t p = 0; p < pos_.size(); ++p) {
    if (pd_body) pd_body[p] = body_[p];
    if (pd_bc) pd_bc[p] = bc_[p];
    if (pd_bc_velocity) pd_bc_velocity[p] = vbc_[p];
  }
}
