// spherocyl_strain: Lees-Edwards shear from a random start or a saved
// configuration, with the argument order of the reference's strain driver
// (src/cpp/main.cpp:63-196):
//   1 data file ("r" for random)   2 output directory   3 N   4 strain rate
//   5 step   6 run length (gamma)   7 position save rate   8 stress save rate
//   9 cell padding   [random only: 10 packing fraction  11 A  12 R
//   13 configuration (0 random, 1 random aligned, 2 zero energy,
//   3 zero energy aligned, 4 grid, 5 grid aligned)]
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
    double rate = float_input(argc, argv, ++argn, "Strain rate");
    double step = float_input(argc, argv, ++argn, "Integration step size");
    double run = float_input(argc, argv, ++argn, "Strain run length (gamma)");
    double pos_rate = float_input(argc, argv, ++argn, "Position data save rate");
    double se_rate = float_input(argc, argv, ++argn, "Stress data save rate");
    double pad = float_input(argc, argv, ++argn, "Cell padding");
    if (se_rate < rate * step) se_rate = rate * step;
    if (pos_rate < rate) pos_rate = rate;

    unsigned long t0 = 0;
    SpherocylBox* box;
    if (file == "r") {
      double phi = float_input(argc, argv, ++argn, "Packing Fraction");
      double A = float_input(argc, argv, ++argn, "Half-shaft length");
      double R = float_input(argc, argv, ++argn, "Radius");
      initialConfig config = RANDOM;
      if (argc > ++argn)
        config = (initialConfig)int_input(argc, argv, argn, "Initial configuration");
      double L = box_side(n, A, R, phi);
      std::cout << "Box length L: " << L << std::endl;
      box = new SpherocylBox(L, n, pad);
      box->set_seed(seed_from_env());
      box->random_configuration(A, R, config);
    } else {
      PositionFile f = read_position_file(file);
      if (f.n != n || (int)f.x.size() != n)
        std::cerr << "Warning: Number of particles in data file may not match requested number"
                  << std::endl;
      if (f.strain_rate != rate)
        std::cerr << "Warning: Strain rate in data file does not match the requested rate"
                  << std::endl;
      if (f.step != step)
        std::cerr << "Warning: Integration step size in data file does not match requested step"
                  << std::endl;
      t0 = position_file_time(file);
      int m = (int)f.x.size();
      box = new SpherocylBox(f.L, m, f.x.data(), f.y.data(), f.phi.data(), f.A.data(), f.R.data(),
                             pad);
      box->set_gamma(f.gamma);
      box->set_total_gamma(f.gamma_total);
    }
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box->set_precision(true);
    std::time_t start = std::time(nullptr);
    box->calc_se();
    box->display(0, 0);
    box->set_output_directory(dir);
    box->set_se_file("sp_strain_se.dat");
    box->set_strain_rate(rate);
    box->set_step(step);
    if (env_flag("SPHEROCYL_APPLY_SAVE_RATES", "1")) {
      box->set_pos_save_step(pos_rate);
      box->set_se_save_step(se_rate);
    }
    box->run_strain(run, t0);
    std::cout << "\nRun Time: " << std::time(nullptr) - start << std::endl;
    delete box;
  } catch (const SpherocylError& e) {
    std::cerr << "spherocyl_strain: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
