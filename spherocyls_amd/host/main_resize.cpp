// spherocyl_resize: compression / expansion at a fixed resize rate to a final
// packing fraction, with the argument order of the reference's resize driver
// (src/cpp/main_resize.cpp:63-194):
//   1 data file ("r" for random)   2 output directory   3 N   4 resize rate
//   5 step   6 final packing fraction   7 position save rate
//   8 stress save rate   9 cell padding   [random only: 10 starting packing
//   fraction  11 A  12 R  13 configuration]
// Saved configurations are read in the save_positions layout (x y phi R A
// after a one-line header); the reference's resize driver reads other columns
// (:140-153), which do not match the files its own box writes.
// SPHEROCYL_RESIZE_IMAGE=current takes the predictor's minimum image with the
// current box length (the reference uses the previous sub-step's,
// src/cpp/resize_dynamics.cpp:101).
#include <ctime>
#include <iostream>
#include <vector>

#include "driver_common.h"

int main(int argc, char* argv[]) {
  using namespace driver;
  try {
    int argn = 0;
    std::string file = string_input(argc, argv, ++argn, "Data file ('r' for random)");
    std::string dir = string_input(argc, argv, ++argn, "Output data directory");
    int n = int_input(argc, argv, ++argn, "Number of particles");
    double rate = float_input(argc, argv, ++argn, "Resize rate");
    double step = float_input(argc, argv, ++argn, "Integration step size");
    double final_phi = float_input(argc, argv, ++argn, "Final packing fraction");
    double pos_rate = float_input(argc, argv, ++argn, "Position data save rate");
    double se_rate = float_input(argc, argv, ++argn, "Stress data save rate");
    double pad = float_input(argc, argv, ++argn, "Cell padding");
    if (se_rate < rate * step) se_rate = rate * step;
    if (pos_rate < rate) pos_rate = rate;

    unsigned long t0 = 0;
    SpherocylBox* box;
    if (file == "r") {
      double phi = float_input(argc, argv, ++argn, "Starting Packing Fraction");
      double A = float_input(argc, argv, ++argn, "Half-shaft length");
      double R = float_input(argc, argv, ++argn, "Radius");
      initialConfig config = RANDOM;
      if (argc > ++argn)
        config = (initialConfig)int_input(argc, argv, argn, "Initial configuration");
      double L = box_side(n, A, R, phi);
      box = new SpherocylBox(L, n, pad);
      box->set_seed(seed_from_env());
      box->random_configuration(A, R, config);
    } else {
      PositionFile f = read_position_file(file);
      if (f.n != n || (int)f.x.size() != n)
        std::cerr << "Warning: Number of particles in data file may not match requested number"
                  << std::endl;
      t0 = position_file_time(file);
      int m = (int)f.x.size();
      box = new SpherocylBox(f.L, m, f.x.data(), f.y.data(), f.phi.data(), f.A.data(), f.R.data(),
                             pad);
    }
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box->set_precision(true);
    if (env_flag("SPHEROCYL_RESIZE_IMAGE", "current")) box->set_resize_image(true);
    std::time_t start = std::time(nullptr);
    box->calc_se();
    box->display(0, 0);
    box->set_output_directory(dir);
    box->set_se_file("sp_resize_se.dat");
    box->set_resize_rate(rate);
    box->set_step(step);
    if (env_flag("SPHEROCYL_APPLY_SAVE_RATES", "1")) {
      box->set_pos_save_step(pos_rate);
      box->set_se_save_step(se_rate);
    }
    box->resize_box(final_phi, t0);
    std::cout << "\nRun Time: " << std::time(nullptr) - start << std::endl;
    delete box;
  } catch (const SpherocylError& e) {
    std::cerr << "spherocyl_resize: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
