// loquat.hpp -- LoquatSph, the simulation driver (loquat.h:25-126).
//
// Same public surface as the reference: LoquatSph(name) and
// RunSimulation(argc, argv) with the `loquat <project> [gpu_id]` CLI; a third
// argument resumes from a checkpoint (every run ends by writing
// <project>/restart.ckpt and its .host companion with the time loop's state).  The
// step runs on the GPU through libretainwall.so; between save points the host
// issues whole batches of steps (hipGraph replay) with no per-step syncs.
#pragma once
#include <chrono>
#include <fstream>
#include <memory>
#include <string>

#include "loquat_grid.hpp"
#include "loquat_io.hpp"

class LoquatSph {
 public:
  explicit LoquatSph(std::string name);
  ~LoquatSph();
  void RunSimulation(int argc, char* argv[]);

 private:
  void GPUInfoExibition(int argc, char* argv[]);   // loquat.cu:289-343
  void InitializeParameters();                     // loquat.cu:130-190
  void CopyDataFromDeviceToHost();                 // loquat.cu:237-251
  void Saving(std::ofstream& time_history_ofs);    // loquat.cu:439-482
  void SummaryToCalculation();                     // loquat.cu:485-509
  static std::string TimeToString(double seconds);

  std::string project_name;
  int selected_device = 0;
  std::unique_ptr<LoquatIo> io;
  std::unique_ptr<LoquatGrid> grid;
  lq_ctx* ctx = nullptr;
  lq_params* par_host = nullptr;
  float probe_box[3] = {0, 0, 0};
  double time = 0.0, output_time = 0.0;
  long long step = 0, step_interval = 0;
  std::chrono::steady_clock::time_point t_initial, t_interval;
  double t_saving = 0.0;
};
