// data_io.cpp -- frames and state files (core/dataIO.cu).
#include "data_io.h"

#include <errno.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <sys/types.h>

#include <string>
#include <vector>

#include "simulator.h"

namespace {

void make_dir(const std::string& d) {
  struct stat info;
  if (stat(d.c_str(), &info) != 0 || !(info.st_mode & S_IFDIR)) mkdir(d.c_str(), 0775);
}

}  // namespace

DataIO::DataIO(SimulationParameters& simParams, RunningParameters& runningParams, Simulator& sim)
    : simParams_(simParams), runningParams_(runningParams), sim_(sim) {}

void DataIO::createOutputFolders() {
  const std::string base = runningParams_.saving_path;
  make_dir(base);
  make_dir(base + "/DATA");
  make_dir(base + "/DATA/STATE");
  make_dir(base + "/DATA/PD_POSITION");
  if (runningParams_.dump_velocity) make_dir(base + "/DATA/PD_VELOCITY");
  if (runningParams_.dump_activity) make_dir(base + "/DATA/PD_ACTIVITY");
  if (runningParams_.dump_pd_bond_list_top) {
    make_dir(base + "/DATA/PD_ORIGINAL_BOND_COUNT");
    make_dir(base + "/DATA/PD_BOND_LIST_TOP");
  }
}

bool DataIO::writeRaw(const char* var, const void* data, size_t bytes) {
  char buff[1200];
  snprintf(buff, sizeof(buff), "%s/DATA/%s/frame.%zu", runningParams_.saving_path, var, savedFrame);
  FILE* f = fopen(buff, "wb");
  if (!f) {
    fprintf(stderr, "Could not open file for write: %s\n", buff);
    return false;
  }
  fwrite(data, 1, bytes, f);
  fclose(f);
  return true;
}

void DataIO::newFrame() { ++savedFrame; }

// core/dataIO.cu:78-120; positions are downloaded and scaled in the host
// mirror in place, as MemoryManager::downloadFromDevice does.
void DataIO::getDataAndWrite() {
  if (!writtenSimulationParameter_) {
    createOutputFolders();
    writeSimulationParameter();
    writtenSimulationParameter_ = true;
  }
  const int n = simParams_.num_pd_particle;
  pd_ctx* ctx = sim_.context();
  if (n <= 0 || !ctx) return;
  real4_t* pos = sim_.get_pd_position();
  if (pd_download(ctx, PD_FIELD_POSITION, pos, sizeof(real4_t) * n) != PD_OK) return;
  for (int i = 0; i < n; ++i) {
    pos[i].x *= simParams_.scaleFactor;
    pos[i].y *= simParams_.scaleFactor;
    pos[i].z *= simParams_.scaleFactor;
    pos[i].w *= simParams_.scaleFactor;
  }
  writeRaw("PD_POSITION", pos, sizeof(real4_t) * n);
  if (runningParams_.dump_velocity) {
    real4_t* vel = sim_.get_pd_velocity();
    if (pd_download(ctx, PD_FIELD_VELOCITY, vel, sizeof(real4_t) * n) == PD_OK)
      writeRaw("PD_VELOCITY", vel, sizeof(real4_t) * n);
  }
  if (runningParams_.dump_activity) writeRaw("PD_ACTIVITY", sim_.get_pd_activity(), sizeof(int) * n);
  if (runningParams_.dump_pd_bond_list_top) {
    std::vector<int> top(n);
    if (pd_download(ctx, PD_FIELD_BOND_TOP, top.data(), sizeof(int) * n) == PD_OK)
      writeRaw("PD_BOND_LIST_TOP", top.data(), sizeof(int) * n);
  }
}

// core/dataIO.cu:256-304
void DataIO::writeSimulationParameter() {
  createOutputFolders();
  char buff[1200];
  snprintf(buff, sizeof(buff), "%s/DATA/sim_info.dat", runningParams_.saving_path);
  FILE* f = fopen(buff, "wb");
  if (f) {
    fwrite(&simParams_, 1, sizeof(SimulationParameters), f);
    fclose(f);
  }
  snprintf(buff, sizeof(buff), "%s/DATA/viz_info.dat", runningParams_.saving_path);
  f = fopen(buff, "w");
  if (!f) return;
  const SimulationParameters& s = simParams_;
  fprintf(f, "num_total_particle %d\n", s.num_total_particle);
  fprintf(f, "num_sph_particle %d\n", s.num_sph_particle);
  fprintf(f, "num_pd_particle %d\n", s.num_pd_particle);
  fprintf(f, "sph_kernel_coeff %g\n", s.sph_kernel_coeff);
  fprintf(f, "pd_kernel_coeff %g\n", s.pd_kernel_coeff);
  fprintf(f, "sph_particle_mass %g\n", s.sph_particle_mass);
  fprintf(f, "pd_particle_mass %g\n", s.pd_particle_mass);
  fprintf(f, "sph_sph_viscosity %g\n", s.sph_sph_viscosity);
  fprintf(f, "sph_boundary_viscosity %g\n", s.sph_boundary_viscosity);
  fprintf(f, "sph_pd_slip %g\n", s.sph_pd_slip);
  fprintf(f, "boundary_min_x %g\n", s.boundary_min_x);
  fprintf(f, "boundary_min_y %g\n", s.boundary_min_y);
  fprintf(f, "boundary_min_z %g\n", s.boundary_min_z);
  fprintf(f, "boundary_max_x %g\n", s.boundary_max_x);
  fprintf(f, "boundary_max_y %g\n", s.boundary_max_y);
  fprintf(f, "boundary_max_z %g\n", s.boundary_max_z);
  fprintf(f, "sph_particle_radius %g\n", s.sph_particle_radius);
  fprintf(f, "sph_rest_density %g\n", s.sph_rest_density);
  fprintf(f, "pd_particle_radius %g\n", s.pd_particle_radius);
  fprintf(f, "pd_horizon %g\n", s.pd_horizon);
  fprintf(f, "scaleFactor %g\n", s.scaleFactor);
  fclose(f);
}

void DataIO::readSimulationParameter(SimulationParameters& _simParams) {
  char buff[1200];
  snprintf(buff, sizeof(buff), "%s/DATA/sim_info.dat", runningParams_.saving_path);
  FILE* f = fopen(buff, "rb");
  if (!f) return;
  if (fread(&_simParams, 1, sizeof(SimulationParameters), f) != sizeof(SimulationParameters))
    fprintf(stderr, "short read: %s\n", buff);
  fclose(f);
}

// State file: savedFrame, then x, v, x0, s, s00, tops, bond lists
// (capacity + list per particle), solution -- everything a step reads.
bool DataIO::saveState(size_t _timestep) {
  writeSimulationParameter();
  pd_ctx* ctx = sim_.context();
  const int n = simParams_.num_pd_particle;
  if (!ctx || n <= 0) return false;
  char buff[1200];
  snprintf(buff, sizeof(buff), "%s/DATA/STATE/state.%zu", runningParams_.saving_path, _timestep);
  FILE* f = fopen(buff, "wb");
  if (!f) return false;
  std::vector<real4_t> v4(n);
  std::vector<double> d1(n);
  std::vector<int> i1(n);
  const int32_t fields4[] = {PD_FIELD_POSITION, PD_FIELD_VELOCITY, PD_FIELD_ORIGINAL_POSITION,
                             PD_FIELD_SOLUTION};
  const int32_t fields1[] = {PD_FIELD_STRETCH, PD_FIELD_ORIGINAL_STRETCH};
  bool ok = fwrite(&savedFrame, 1, sizeof(size_t), f) == sizeof(size_t);
  for (int32_t fd : fields4) {
    ok = ok && pd_download(ctx, fd, v4.data(), sizeof(real4_t) * n) == PD_OK;
    ok = ok && fwrite(v4.data(), sizeof(real4_t), n, f) == (size_t)n;
  }
  for (int32_t fd : fields1) {
    ok = ok && pd_download(ctx, fd, d1.data(), sizeof(double) * n) == PD_OK;
    ok = ok && fwrite(d1.data(), sizeof(double), n, f) == (size_t)n;
  }
  ok = ok && pd_download(ctx, PD_FIELD_BOND_TOP, i1.data(), sizeof(int) * n) == PD_OK;
  ok = ok && fwrite(i1.data(), sizeof(int), n, f) == (size_t)n;
  std::vector<int> cap(n), lists((size_t)n * MAX_PD_BOND_COUNT);
  ok = ok && pd_get_bonds(ctx, cap.data(), lists.data(), MAX_PD_BOND_COUNT) == PD_OK;
  ok = ok && fwrite(cap.data(), sizeof(int), n, f) == (size_t)n;
  for (int p = 0; ok && p < n; ++p)
    ok = fwrite(lists.data() + (size_t)p * MAX_PD_BOND_COUNT, sizeof(int), cap[p], f) == (size_t)cap[p];
  fclose(f);
  return ok;
}

// core/dataIO.cu:152-198: the saved SimulationParameters (time step,
// phase-in clock) come back from sim_info.dat after the particle counts are
// checked; the step counter resumes at the saved step.
bool DataIO::loadState(size_t _timestep) {
  pd_ctx* ctx = sim_.context();
  const int n = simParams_.num_pd_particle;
  if (!ctx || n <= 0) return false;
  {
    SimulationParameters saved = simParams_;
    readSimulationParameter(saved);
    if (saved.num_pd_particle != simParams_.num_pd_particle ||
        saved.num_sph_particle != simParams_.num_sph_particle) {
      fprintf(stderr, "Inconsistent number of particles with saved data.\n");
      return false;
    }
    simParams_ = saved;
  }
  writtenSimulationParameter_ = true;
  char buff[1200];
  snprintf(buff, sizeof(buff), "%s/DATA/STATE/state.%zu", runningParams_.saving_path, _timestep);
  FILE* f = fopen(buff, "rb");
  if (!f) return false;
  std::vector<real4_t> v4(n);
  std::vector<double> d1(n);
  std::vector<int> top(n), cap(n), lists((size_t)n * MAX_PD_BOND_COUNT, 0);
  bool ok = fread(&savedFrame, 1, sizeof(size_t), f) == sizeof(size_t);
  const int32_t fields4[] = {PD_FIELD_POSITION, PD_FIELD_VELOCITY, PD_FIELD_ORIGINAL_POSITION,
                             PD_FIELD_SOLUTION};
  for (int32_t fd : fields4) {
    ok = ok && fread(v4.data(), sizeof(real4_t), n, f) == (size_t)n;
    ok = ok && pd_upload(ctx, fd, v4.data(), sizeof(real4_t) * n) == PD_OK;
  }
  const int32_t fields1[] = {PD_FIELD_STRETCH, PD_FIELD_ORIGINAL_STRETCH};
  for (int32_t fd : fields1) {
    ok = ok && fread(d1.data(), sizeof(double), n, f) == (size_t)n;
    ok = ok && pd_upload(ctx, fd, d1.data(), sizeof(double) * n) == PD_OK;
  }
  ok = ok && fread(top.data(), sizeof(int), n, f) == (size_t)n;
  ok = ok && fread(cap.data(), sizeof(int), n, f) == (size_t)n;
  for (int p = 0; ok && p < n; ++p) {
    if (cap[p] < 0 || cap[p] > MAX_PD_BOND_COUNT) ok = false;
    else
      ok = fread(lists.data() + (size_t)p * MAX_PD_BOND_COUNT, sizeof(int), cap[p], f) ==
           (size_t)cap[p];
  }
  fclose(f);
  if (!ok) return false;
  // restore the full lists, then the live counts
  if (pd_set_bonds(ctx, cap.data(), lists.data(), MAX_PD_BOND_COUNT) != PD_OK) return false;
  if (pd_upload(ctx, PD_FIELD_BOND_TOP, top.data(), sizeof(int) * n) != PD_OK) return false;
  sim_.resumeAt((int)_timestep);
  return true;
}
