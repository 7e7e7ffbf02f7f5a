// scene.cpp -- mesh-filled scenes, box-lattice specimens and a seeded
// spherical projectile.
//
// Mesh (examples/scene.cpp:18-136, 360-395, 530-598): the OBJ mesh is
// scaled to half of 0.9 x the smallest extent of the boundary box over its
// largest extent, its bounding box moved to the boundary minimum, and the
// lattice of spacing 2r over that box (x fastest) keeps the points inside
// the mesh, clamped to the box; then the particles are rotated and moved
// by mesh_translation / scaleFactor.  The float roundings of the reference
// (float mesh coordinates and box, fminf / fmaxf) are kept.
//
// Specimen: nx x ny x nz simple-cubic lattice, x fastest (the reference's
// fill order, examples/scene.cpp:536-540), spacing h, each coordinate
// displaced by (2u - 1) * jitter * h with u = scene_uniform(seed, 3 p + axis).
// Dirichlet layers: z-layers k < bc_fix_zmin_layers are held at v = 0,
// k >= nz - bc_move_zmax_layers move at (0, 0, bc_move_zmax_velocity).
//
// Batch (scene batch): batch_count independent tension specimens.  Specimen
// k draws, from scene_uniform(batch_seed, 3k + {0, 1, 2}), its particle count
// log-uniformly in [batch_min, batch_max], its aspect a uniformly in
// [1, batch_max_aspect] (nx = ny = round((N/a)^(1/3)), nz = round(N/nx^2),
// the long axis along the pull) and its s0 uniformly in [s0_min, s0_max].
// Specimens stand on an xy grid, one horizon plus the jitter apart at least;
// each is jittered like the lattice (index (k << 40) + 3p + axis) and loaded
// like it (bc_* layers).  Rows of specimen k start on a multiple of 256;
// the rows between specimens are INVISIBLE copies of the specimen's first
// particle.
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

#include "mesh_fill.h"

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
    if (simParams_.num_pd_particle == 0) {  // examples/scene.cpp:25-28
      ok_ = true;
      return;
    }
    if (!buildMesh()) return;
    simParams_.num_pd_particle = (int)pos_.size();
    simParams_.num_total_particle = simParams_.num_sph_particle + simParams_.num_pd_particle;
    simParams_.num_clists = simParams_.num_pd_particle;
    while (simParams_.num_clists % 8 != 0) simParams_.num_clists++;
    ok_ = true;
    return;
  }
  const RunningParameters& r = runningParams_;
  if (r.scene_lattice == 2) {
    if (!buildBatch()) return;
    simParams_.num_pd_particle = (int)pos_.size();
    simParams_.num_total_particle = simParams_.num_sph_particle + simParams_.num_pd_particle;
    simParams_.num_clists = simParams_.num_pd_particle;
    while (simParams_.num_clists % 8 != 0) simParams_.num_clists++;
    ok_ = true;
    return;
  }
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
    if (pd_activity) pd_activity[p] = act_.empty() ? ACTIVE : act_[p];
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

// examples/scene.cpp:18-136 (scaling, grid) and 530-598 (fill).
bool Scene::buildMesh() {
  const SimulationParameters& sp = simParams_;
  ObjMesh mesh;
  std::string e;
  if (!load_obj(runningParams_.obj_file, mesh, e)) {
    err_msg_ = e;
    err_ = err_msg_.c_str();
    return false;
  }
  const size_t nv = mesh.v.size() / 3;
  // cyTriMesh::ComputeBoundingBox over all vertices, in float
  float bmin[3] = {mesh.v[0], mesh.v[1], mesh.v[2]};
  float bmax[3] = {mesh.v[0], mesh.v[1], mesh.v[2]};
  for (size_t i = 1; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float c = mesh.v[3 * i + a];
      if (bmin[a] > c) bmin[a] = c;
      if (bmax[a] < c) bmax[a] = c;
    }
  float dif[3];
  for (int a = 0; a < 3; ++a) dif[a] = bmax[a] - bmin[a];
  const double maxDiff = (double)fmaxf(fmaxf(fabsf(dif[0]), fabsf(dif[1])), fabsf(dif[2]));
  const float ext = fminf(fminf((float)(sp.boundary_max_x - sp.boundary_min_x),
                                (float)(sp.boundary_max_y - sp.boundary_min_y)),
                          (float)(sp.boundary_max_z - sp.boundary_min_z));
  double scale = (double)ext * 0.9 / maxDiff;
  scale *= 0.5f;
  const double bnd_min[3] = {sp.boundary_min_x, sp.boundary_min_y, sp.boundary_min_z};
  double shift[3];
  for (int a = 0; a < 3; ++a) {
    bmin[a] = (float)((double)bmin[a] * scale);
    bmax[a] = (float)((double)bmax[a] * scale);
    shift[a] = bnd_min[a] - (double)bmin[a];
    bmin[a] = (float)((double)bmin[a] + shift[a]);
    bmax[a] = (float)((double)bmax[a] + shift[a]);
  }
  printf("    Bounding box for mesh object: [%g, %g, %g] -> [%g, %g, %g]\n", bmin[0], bmin[1],
         bmin[2], bmax[0], bmax[1], bmax[2]);
  const double r = sp.pd_particle_radius;
  int grid[3];
  for (int a = 0; a < 3; ++a) {
    const float span = bmax[a] - bmin[a];
    grid[a] = (int)ceil((double)span / (2.0f * r)) + 1;
  }
  printf("    Maximum Peridynamics grid: %dx%dx%d\n", grid[0], grid[1], grid[2]);
  mesh_xyz_.resize(3 * nv);
  for (size_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) mesh_xyz_[3 * i + a] = (double)mesh.v[3 * i + a] * scale + shift[a];
  mesh_tri_ = mesh.f;
  MeshInside inside(mesh_xyz_, mesh_tri_);
  const double jitter = arrangement_ == REGULAR_GRID ? 0.0 : kMeshJitter * r;
  const double spacing = 2 * r;
  pos_.clear();
  for (int z = 0; z < grid[2]; ++z)
    for (int y = 0; y < grid[1]; ++y)
      for (int x = 0; x < grid[0]; ++x) {
        // three rand() draws per lattice point, as the reference's frand()
        double p[3];
        const int idx[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
          const double u = rand() / (double)RAND_MAX;
          p[a] = (double)bmin[a] + idx[a] * spacing + (u * 2.0f - 1.0f) * jitter;
        }
        if (!inside.inside(p)) continue;
        for (int a = 0; a < 3; ++a) {
          p[a] = (p[a] < (double)bmin[a]) ? (double)bmin[a] : p[a];
          p[a] = (p[a] > (double)bmax[a]) ? (double)bmax[a] : p[a];
        }
        pos_.push_back(make_real4(p[0], p[1], p[2], 0.0));
      }
  printf("    Total filled Peridynamics particles: %zu\n", pos_.size());
  // examples/scene.cpp:360-395: rotation (zero angles) then translation
  const double tr[3] = {(double)runningParams_.mesh_translation_x / sp.scaleFactor,
                        (double)runningParams_.mesh_translation_y / sp.scaleFactor,
                        (double)runningParams_.mesh_translation_z / sp.scaleFactor};
  const double az = 0.0, el = 0.0, ro = 0.0;
  const double sA = sin(az), cA = cos(az), sE = sin(el), cE = cos(el), sR = sin(ro), cR = cos(ro);
  const double R[3][3] = {{cR * cA - sR * sA * sE, sR * cA + cR * sA * sE, -sA * cE},
                          {-sR * cE, cR * cE, sE},
                          {cR * sA + sR * cA * sE, sR * sA - cR * cA * sE, cA * cE}};
  const double v0 = (double)runningParams_.pd_initial_velocity;
  vel_.assign(pos_.size(), make_real4(v0, v0, v0, v0));
  for (real4_t& q : pos_) {
    double t[3];
    for (int j = 0; j < 3; ++j) t[j] = R[j][0] * q.x + R[j][1] * q.y + R[j][2] * q.z;
    q = make_real4(t[0] + tr[0], t[1] + tr[1], t[2] + tr[2], 0.0);
  }
  body_.assign(pos_.size(), 0);
  bc_.assign(pos_.size(), 0);
  vbc_.assign(pos_.size(), make_real4(0, 0, 0, 0));
  return true;
}

bool Scene::buildBatch() {
  const RunningParameters& r = runningParams_;
  const int nspec = r.batch_count;
  const double h = r.lattice_spacing;
  if (nspec < 1 || !(h > 0.0) || r.batch_min_particles < 8 ||
      r.batch_max_particles < r.batch_min_particles || !(r.batch_max_aspect >= 1.0)) {
    err_ = "batch_count, lattice_spacing, batch_min/max_particles and batch_max_aspect are invalid";
    return false;
  }
  const double jit = r.lattice_jitter;
  struct Spec {
    int nx, ny, nz;
    float s0;
  };
  std::vector<Spec> spec(nspec);
  int max_nx = 0;
  for (int k = 0; k < nspec; ++k) {
    const double u0 = scene_uniform(r.batch_seed, 3ull * k);
    const double u1 = scene_uniform(r.batch_seed, 3ull * k + 1);
    const double u2 = scene_uniform(r.batch_seed, 3ull * k + 2);
    const double lo = log((double)r.batch_min_particles), hi = log((double)r.batch_max_particles);
    const double N = exp(lo + u0 * (hi - lo));
    const double a = 1.0 + u1 * (r.batch_max_aspect - 1.0);
    Spec& s = spec[k];
    s.nx = s.ny = std::max(2, (int)llround(cbrt(N / a)));
    s.nz = std::max(2, (int)llround(N / ((double)s.nx * s.ny)));
    s.s0 = (float)(r.batch_s0_min + u2 * (r.batch_s0_max - r.batch_s0_min));
    max_nx = std::max(max_nx, s.nx);
  }
  const int cols = (int)ceil(sqrt((double)nspec));
  const double gap = simParams_.pd_horizon + 2.0 * jit * h + h;
  const double pitch = (max_nx - 1) * h + gap;
  pos_.clear();
  vel_.clear();
  body_.clear();
  bc_.clear();
  vbc_.clear();
  act_.clear();
  s0_.clear();
  spec_first_.clear();
  spec_count_.clear();
  for (int k = 0; k < nspec; ++k) {
    const Spec& s = spec[k];
    // pad to the next 256-row boundary with INVISIBLE copies of the previous
    // specimen's first particle
    while (pos_.size() % 256 != 0) {
      const size_t src = spec_first_.back();
      pos_.push_back(pos_[src]);
      vel_.push_back(make_real4(0, 0, 0, 0));
      body_.push_back(0);
      bc_.push_back(0);
      vbc_.push_back(make_real4(0, 0, 0, 0));
      act_.push_back(INVISIBLE);
      s0_.push_back(s0_.back());
    }
    spec_first_.push_back((int)pos_.size());
    spec_count_.push_back(s.nx * s.ny * s.nz);
    const double org[3] = {(k % cols) * pitch, (k / cols) * pitch, 0.0};
    for (int kz = 0; kz < s.nz; ++kz)
      for (int j = 0; j < s.ny; ++j)
        for (int i = 0; i < s.nx; ++i) {
          const uint64_t p = ((uint64_t)kz * s.ny + j) * s.nx + i;
          const int idx[3] = {i, j, kz};
          double c[3];
          for (int a = 0; a < 3; ++a) {
            const double u = scene_uniform(r.lattice_seed, ((uint64_t)k << 40) + 3 * p + a);
            c[a] = org[a] + idx[a] * h + (2.0 * u - 1.0) * jit * h;
          }
          pos_.push_back(make_real4(c[0], c[1], c[2], 0.0));
          real4_t v = make_real4(0, 0, 0, 0), vb = v;
          int bc = 0;
          if (kz < r.bc_fix_zmin_layers) {
            bc = 1;
          } else if (kz >= s.nz - r.bc_move_zmax_layers) {
            bc = 1;
            vb = make_real4(0.0, 0.0, r.bc_move_zmax_velocity, 0.0);
            v = vb;
          }
          vel_.push_back(v);
          body_.push_back(0);
          bc_.push_back(bc);
          vbc_.push_back(vb);
          act_.push_back(ACTIVE);
          s0_.push_back(s.s0);
        }
  }
  return true;
}
