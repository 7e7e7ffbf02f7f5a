// data_io.h -- frame and state files of the reference (core/dataIO.h) so that
// its viewer reads this engine's output: raw real4 frames under
// <saving_path>/DATA/<VARIABLE>/frame.<n>, positions multiplied by the scale
// factor, plus DATA/sim_info.dat (raw SimulationParameters) and
// DATA/viz_info.dat (text).
#ifndef PD_HOST_DATA_IO_H
#define PD_HOST_DATA_IO_H

#include <stddef.h>

#include "parameters.h"

class Simulator;

class DataIO {
 public:
  DataIO(SimulationParameters& simParams, RunningParameters& runningParams, Simulator& sim);

  void readSimulationParameter(SimulationParameters& _simParams);
  void writeSimulationParameter();
  // core/dataIO.cu:78-120: download and write every selected variable
  void getDataAndWrite();
  // core/dataIO.cu:123-198 with the state file under DATA/STATE (the folder
  // the reference creates; it wrote one level up)
  bool saveState(size_t _timestep);
  bool loadState(size_t _timestep);
  void newFrame();

  size_t savedFrame = 0;

 private:
  void createOutputFolders();
  bool writeRaw(const char* var, const void* data, size_t bytes);

  SimulationParameters& simParams_;
  RunningParameters& runningParams_;
  Simulator& sim_;
  bool writtenSimulationParameter_ = false;
};

#endif
