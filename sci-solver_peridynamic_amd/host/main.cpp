// main.cpp -- peri_run: the reference's example driver (examples/main.cpp)
// on this engine.  Usage: peri_run <params.txt> [obj_file] [out_dir]
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "parameters.h"
#include "scene.h"
#include "simulator.h"

int main(int argc, char* argv[]) {
  if (argc < 2) {
    printf("Usage: ./peri_run <parameter_file> [obj_file] [out_dir]\n");
    return EXIT_FAILURE;
  }
  SystemParameter sysParams(argv[1]);
  if (!sysParams.ok()) return EXIT_FAILURE;
  RunningParameters& runningParams = sysParams.getRunningParameters();
  if (argc > 2) snprintf(runningParams.obj_file, sizeof(runningParams.obj_file), "%s", argv[2]);
  if (argc > 3) snprintf(runningParams.saving_path, sizeof(runningParams.saving_path), "%s", argv[3]);
  SimulationParameters& simParams = sysParams.getSimulationParameters();
  Scene scene(simParams, runningParams, Scene::REGULAR_GRID);
  if (!scene.ok()) {
    fprintf(stderr, "%s\n", scene.error());
    return EXIT_FAILURE;
  }
  Simulator simulator(runningParams.gpu_device, simParams, runningParams);
  if (!simulator.ok()) return EXIT_FAILURE;
  scene.initPeridynamicsParticles(simulator.get_pd_activity(), simulator.get_pd_position(),
                                  simulator.get_pd_velocity());
  scene.initPeridynamicsBoundary(simulator.get_pd_body(), simulator.get_pd_bc(),
                                 simulator.get_pd_bc_velocity());
  if (scene.numSpecimens() > 0) {
    simulator.setSpecimens(scene.specimenFirst(), scene.specimenCount());
    simulator.setStretchLimits(scene.particleS0());
  }
  simulator.makeReady();
  if (runningParams.start_step > 0 && !simulator.dataIO().loadState(runningParams.start_step)) {
    fprintf(stderr, "could not load state %d\n", runningParams.start_step);
    return EXIT_FAILURE;
  }
  if (!simulator.ok()) return EXIT_FAILURE;
  const bool write = runningParams.saving_path[0] != 0;
  int frameStep = 0, stateStep = 0;
  while (!simulator.finished()) {
    if (frameStep == 0 && write) {
      simulator.dataIO().newFrame();
      simulator.dataIO().getDataAndWrite();
      simulator.printSimulationTime();
    }
    if (write && runningParams.step_per_state > 0 && stateStep == runningParams.step_per_state) {
      simulator.dataIO().saveState(simulator.get_current_step());
      stateStep = 0;
    }
    simulator.advance();
    ++frameStep;
    ++stateStep;
    if (frameStep == runningParams.step_per_frame) frameStep = 0;
  }
  if (write) {
    simulator.dataIO().newFrame();
    simulator.dataIO().getDataAndWrite();
  }
  simulator.printSimulationTime();
  printf("Total simulation steps: %d\n", runningParams.final_step);
  if (!simulator.ok()) {
    printf("Simulation failed: %s\n", simulator.error().c_str());
    return EXIT_FAILURE;
  }
  if (simulator.CGConvergence()) printf("Simulation finished successfully.\n");
  else printf("Simulation failed.\n");
  return simulator.CGConvergence() ? EXIT_SUCCESS : 2;
}
