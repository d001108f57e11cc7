// spherocyl_minimize: gradient-descent relaxation with the argument order of
// the reference's relaxation driver (src/cpp/main_relax.cpp:63-182):
//   1 data file ("r" for random)   2 output directory   3 N   4 step
//   5 maximum relaxation steps   6 position save rate (steps)
//   7 stress save rate (steps)   8 cell padding
//   [random only: 9 packing fraction  10 A  11 R]
// SPHEROCYL_MINIMIZER=fire replaces gradient descent by FIRE to F_rms < 1e-10
// (positions written as sr<iterations>.dat).
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
    double step = float_input(argc, argv, ++argn, "Integration step size");
    int nmax = int_input(argc, argv, ++argn, "Maximum relaxation steps (n)");
    int pos_rate = int_input(argc, argv, ++argn, "Position data save rate (n)");
    int se_rate = int_input(argc, argv, ++argn, "Stress data save rate (n)");
    (void)se_rate;  // parsed but unused, as in the reference
    double pad = float_input(argc, argv, ++argn, "Cell padding");

    SpherocylBox* box;
    if (file == "r") {
      double phi = float_input(argc, argv, ++argn, "Packing Fraction");
      double A = float_input(argc, argv, ++argn, "Half-shaft length");
      double R = float_input(argc, argv, ++argn, "Radius");
      double L = box_side(n, A, R, phi);
      box = new SpherocylBox(L, n, pad);
      box->set_seed(seed_from_env());
      box->random_configuration(A, R);
    } else {
      PositionFile f = read_position_file(file);
      if (f.n != n || (int)f.x.size() != n)
        std::cerr << "Warning: Number of particles in data file may not match requested number"
                  << std::endl;
      if (f.step != step)
        std::cerr << "Warning: Integration step size in data file does not match requested step"
                  << std::endl;
      int m = (int)f.x.size();
      box = new SpherocylBox(f.L, m, f.x.data(), f.y.data(), f.phi.data(), f.A.data(), f.R.data(),
                             pad);
    }
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box->set_precision(true);
    std::time_t start = std::time(nullptr);
    box->calc_se();
    box->display(0, 0);
    box->set_output_directory(dir);
    box->set_se_file("sr_strain_se.dat");
    box->set_strain_rate(0);
    box->set_step(step);
    box->set_pos_save_step((double)pos_rate);
    if (env_flag("SPHEROCYL_MINIMIZER", "fire")) {
      int it = box->fire_minimize();
      char name[32];
      std::snprintf(name, sizeof name, "sr%010d.dat", it);
      box->save_positions(name);
      std::cout << "FIRE iterations: " << it << "  F_rms: " << box->residual_force() << std::endl;
    } else {
      box->gradient_descent_minimize(nmax, 0.0);
    }
    box->display(0, 0);
    std::cout << "\nRun Time: " << std::time(nullptr) - start << std::endl;
    delete box;
  } catch (const SpherocylError& e) {
    std::cerr << "spherocyl_minimize: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
