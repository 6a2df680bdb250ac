// loquat.cpp -- LoquatSph driver over libretainwall.so.
#include "loquat.hpp"

#include <sys/stat.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <vector>

LoquatSph::LoquatSph(std::string name) : project_name(std::move(name)) {
  io = std::make_unique<LoquatIo>(project_name);
}

LoquatSph::~LoquatSph() {
  if (ctx) lq_destroy(ctx);
}

std::string LoquatSph::TimeToString(double seconds) {  // loquat.cu:106-127
  long long s = (long long)seconds;
  std::ostringstream o;
  o << std::setfill('0') << std::setw(2) << s / 86400 << "d " << std::setw(2) << (s % 86400) / 3600 << "h "
    << std::setw(2) << (s % 3600) / 60 << "m " << std::setw(2) << s % 60 << "s ";
  return o.str();
}

void LoquatSph::GPUInfoExibition(int argc, char* argv[]) {
  std::cout << "**********************************" << std::endl;
  std::cout << "** LOQUAT Solver for single GPU **" << std::endl;
  std::cout << "**********************************" << std::endl << std::endl;
  if (argc < 2) {
    std::cout << "Error! Please give the project name!" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  selected_device = argc >= 3 ? std::atoi(argv[2]) : 0;
  struct stat st;
  if (stat(project_name.c_str(), &st) != 0) {
    std::cout << "Create results folder: " << project_name << std::endl;
    mkdir(project_name.c_str(), 0755);
  }
  std::cout << "Simulation running on GPU " << selected_device << std::endl;
}

void LoquatSph::InitializeParameters() {
  par_host->interaction_variant = LQ_INTERACTION_NEIGHBOR_LIST;  // fastest bit-identical variant
  if (lq_initialize_parameters(par_host) != LQ_OK) {
    std::cout << "The input dimension of problem is " << par_host->dim << std::endl;
    std::cout << "Check the dimension of problem!" << std::endl;
    std::exit(EXIT_FAILURE);
  }
  std::printf("ncellx = %d, ncelly = %d, ncellz = %d, ncell = %d\n", par_host->cell_num[0], par_host->cell_num[1],
              par_host->cell_num[2], par_host->ncell);
  std::printf("ncell_material = %d  \nncell_delete_m = %d  \nncell = %d  \n\n", par_host->ncell_material,
              par_host->ncell_delete_m, par_host->ncell);
  // wall-force probe of utilities.cu:619-621 (case-1 constants)
  const float bt = par_host->bound_thickness;
  probe_box[0] = 0.04f + bt + par_host->dr * 3;
  probe_box[1] = bt + 0.04f;
  probe_box[2] = bt + 0.04f + 1.0f;
}

void LoquatSph::CopyDataFromDeviceToHost() {
  const size_t np = size_t(par_host->np);
  int32_t npb = 0, npm = 0, ndel = 0;
  lq_check(ctx, lq_get_counts(ctx, &npb, &npm, &ndel), "lq_get_counts");
  par_host->npm = npm;
  par_host->np_delete_m = ndel;
  lq_check(ctx, lq_get_field(ctx, LQ_F_CODES, io->GetCodes().data(), np * 16), "codes");
  lq_check(ctx, lq_positions_absolute(ctx, io->GetPosPres().data(), np * 16), "positions");
  lq_check(ctx, lq_get_field(ctx, LQ_F_VEL, io->GetVelRhop().data(), np * 16), "velrhop");
  lq_check(ctx, lq_get_field(ctx, LQ_F_IDM, io->GetIdMassIdx().data(), np * 16), "idmassidx");
  std::vector<float> nrm(np * 4);
  lq_check(ctx, lq_get_field(ctx, LQ_F_NORMAL, nrm.data(), np * 16), "normal");
  std::copy(nrm.begin(), nrm.begin() + size_t(par_host->npb) * 4, io->GetNormalMaterial().begin());
  lq_check(ctx, lq_stress_saving_regularization(ctx, io->GetStr().data(), np * 24), "StressSavingRegularization");
  lq_check(ctx, lq_get_field(ctx, LQ_F_STN, io->GetStn().data(), np * 24), "stn");
  lq_check(ctx, lq_get_field(ctx, LQ_F_ACCAD, io->GetAccad().data(), np * 16), "accad");
}

void LoquatSph::Saving(std::ofstream& th) {
  auto t0 = std::chrono::steady_clock::now();
  double interval = std::chrono::duration<double>(t0 - t_interval).count();
  double fps = double(step - step_interval + 1) / std::max(interval, 1e-9);
  double used = std::chrono::duration<double>(t0 - t_initial).count();
  std::printf(" \n----------Save-----------\n");
  std::printf("Time = %0.5f | Step = %lld | FPS = %0.5f\n", par_host->time, step, fps);
  std::cout << "Clock time used  : " << TimeToString(used) << std::endl;
  std::printf("Number of particles: %d | Number of cells: %d\n", par_host->np, par_host->ncell);
  par_host->output_number++;
  CopyDataFromDeviceToHost();
  io->SaveXml();
  io->SaveParticles();
  {
    // GetValuesAlongLine (utilities.cu:324-367) with the reference's case
    // constants: wall face 0.04 + bt - 0.005 t, depths [bt + 0.04, bt + 1.04]
    const float bt = par_host->bound_thickness;
    const int nip = int(1.0f / par_host->dr);
    const float geom[4] = {float(0.04f + bt - double(par_host->time) * 0.005), bt + 0.04f, bt + 0.04f + 1.0f,
                           0.04f + 1.0f + bt};
    std::vector<float> prof(size_t(std::max(nip, 1)) * 4);
    if (nip > 0)
      lq_check(ctx, lq_wall_pressure_profile(ctx, geom, nip, prof.data(), prof.size() * sizeof(float)),
               "GetValuesAlongLine");
    std::ofstream zz(project_name + "/stress_zz" + std::to_string(par_host->output_number) + ".txt");
    for (int i = 0; i < nip; i++)
      zz << prof[4 * size_t(i)] << " " << prof[4 * size_t(i) + 1] << " " << prof[4 * size_t(i) + 2] << " "
         << prof[4 * size_t(i) + 3] << std::endl;
  }
  double F = 0.0;
  lq_check(ctx, lq_wall_force(ctx, probe_box, &F), "GetValuesAtPoint");
  th << par_host->time << " " << F << std::endl;  // utilities.cu:669
  std::cout << "Save file: " << project_name << "/" << par_host->output_number << ".dat + output_"
            << par_host->output_number << ".xml" << std::endl << std::endl;
  output_time = 0.0;
  step_interval = step;
  t_interval = std::chrono::steady_clock::now();
  t_saving += std::chrono::duration<double>(t_interval - t0).count();
}

void LoquatSph::SummaryToCalculation() {
  double total = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_initial).count();
  double fps = double(step) / std::max(total - t_saving, 1e-9);
  std::cout << "Simulation finished ..." << std::endl << std::endl << "----------SUMMARY----------" << std::endl;
  std::cout << "1. Total clock time = " << TimeToString(total) << std::endl;
  std::cout << "7. Saving = " << TimeToString(t_saving) << std::endl;
  std::cout << "8. Total steps = " << step << std::endl;
  std::cout << "9. Average time step = " << (step ? par_host->time / float(step) : 0.f) << std::endl;
  std::cout << "10. Average FPS = " << fps << std::endl;
}

void LoquatSph::RunSimulation(int argc, char* argv[]) {
  GPUInfoExibition(argc, argv);
  io->LoadParameters();
  io->LoadParticles();
  par_host = io->GetParameters();
  InitializeParameters();
  lq_check(nullptr, lq_create(par_host, selected_device, &ctx), "lq_create");
  grid = std::make_unique<LoquatGrid>(ctx);
  lq_host_state hs{io->GetCodes().data(), io->GetPosPres().data(), io->GetVelRhop().data(),
                   io->GetIdMassIdx().data(), io->GetStr().data(), io->GetStn().data(),
                   io->GetNormalMaterial().data()};
  lq_check(ctx, lq_upload(ctx, &hs), "CopyDataFromHostToDevice");
  if (par_host->n_motion > 0)
    lq_check(ctx, lq_set_motion_tables(ctx, par_host->n_motion, io->GetMotionTables().data()), "motion tables");
  grid->GridConfig(par_host);
  grid->NeighborSearch();
  lq_check(ctx, lq_sort_arrays_primary(ctx), "SortArraysPrimary");
  const bool resume = argc >= 4;
  t_initial = t_interval = std::chrono::steady_clock::now();
  step_interval = 0;
  if (resume) {
    // loquat <project> <gpu> <checkpoint>: the device state from the
    // checkpoint, the time loop's own state from its .host companion
    const std::string ck = argv[3];
    lq_check(ctx, lq_checkpoint_read(ctx, ck.c_str()), "lq_checkpoint_read");
    std::ifstream hin(ck + ".host");
    int out_no = 0;
    if (!(hin >> time >> output_time >> step >> out_no)) {
      std::cout << "Error! Cannot read " << ck << ".host !" << std::endl;
      std::exit(EXIT_FAILURE);
    }
    par_host->output_number = out_no;
    par_host->time = float(time);
    step_interval = step;
    std::cout << "Resuming from " << ck << " at time " << time << ", step " << step << std::endl;
  } else {
    std::printf("Saving initial state ...\n");
    CopyDataFromDeviceToHost();
    io->SaveXml();
    io->SaveParticles();
    time = double(par_host->time);
    step = 0;
    output_time = 0.0;
    lq_check(ctx, lq_set_time(ctx, time, 0), "lq_set_time");
  }
  std::ofstream th(project_name + "/time_history.txt", resume ? std::ios::app : std::ios::trunc);
  // dt = cfl * h / cs in float, as StepCalculation (loquat.cu:368-369)
  const float dt = par_host->cfl * (par_host->h / par_host->cs_m);
  const double tmax = double(par_host->time_max), freq = double(par_host->output_frequency);
  while (time < tmax) {
    // run up to (and including) the next step whose end triggers a save
    long long n = 0;
    bool save = false;
    double t = time, ot = output_time;
    while (t < tmax) {
      n++;
      if (ot > freq) { save = true; break; }
      t += double(dt);
      ot += double(dt);
    }
    lq_check(ctx, lq_step(ctx, int(n), 1), "step");
    if (save) {
      time = t;
      output_time = ot;
      step += n - 1;
      par_host->time = float(time);
      par_host->dt = dt;
      Saving(th);
      time += double(dt);
      output_time += double(dt);
      step += 1;
    } else {
      time = t;
      output_time = ot;
      step += n;
    }
  }
  lq_check(ctx, lq_synchronize(ctx), "synchronize");
  par_host->time = float(time);
  // restart point: a later run with a larger time-max continues from here
  const std::string ck = project_name + "/restart.ckpt";
  lq_check(ctx, lq_checkpoint_write(ctx, ck.c_str()), "lq_checkpoint_write");
  std::ofstream hout(ck + ".host");
  hout.precision(17);
  hout << time << " " << output_time << " " << step << " " << par_host->output_number << std::endl;
  SummaryToCalculation();
}
