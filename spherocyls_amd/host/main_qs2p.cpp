// spherocyl_qscompress / spherocyl_findjam: the quasi-static protocols with the
// argument order of the reference's GPU drivers (src/cuda/main_2point_qscompress.cpp
// and main_2point_findjam.cpp, lines 60-251):
//   1 data file ("r" for random)   2 output directory   3 N
//   4 resize rate (epsilon)   5 integration step size
//   6 final packing fraction (qscompress) | energy at jamming (findjam)
//   7 position save rate   8 stress save rate   9 cell padding
//   [random only: 10 packing fraction  11 aspect ratio A/R  12 bidispersity
//    13 initial configuration (0 random, 2 random-aligned, 3 zero-energy,
//       5 zero-energy-aligned, 6 grid, 8 grid-aligned)]
// Random starts: R = 0.5 and A = aspect R for even indices, both times the
// bidispersity for odd ones (src/cuda/main_2point_qscompress.cpp:164-170).
// A data file is either format: save_positions (sp*.dat, "N L ..." header)
// or the GPU code's sd*.dat ("N Lx Ly ..." header; Lx must equal Ly here).
// Contacts use the two-point model; minimisation is FIRE to F_rms < 1e-10.
// SPHEROCYL_PRECISION=mixed selects FP32 pair forces; SPHEROCYL_SEED seeds
// random starts.
#include <ctime>
#include <iostream>
#include <vector>

#include "driver_common.h"
#include "spherocyl_files.h"

namespace {

int config_from_gpu_code(int c) {
  switch (c) {
    case 0: return SC_CONFIG_RANDOM;
    case 2: return SC_CONFIG_RANDOM_ALIGNED;
    case 3: return SC_CONFIG_ZERO_E;
    case 5: return SC_CONFIG_ZERO_E_ALIGNED;
    case 6: return SC_CONFIG_GRID;
    case 8: return SC_CONFIG_GRID_ALIGNED;
    default:
      throw SpherocylError(SC_ERR_ARG, "initial configuration " + std::to_string(c) +
                                           " is not supported (0, 2, 3, 5, 6, 8 are)");
  }
}

}  // namespace

int main(int argc, char* argv[]) {
  using namespace driver;
#ifdef SPHEROCYL_FIND_JAM
  const char* self = "spherocyl_findjam";
#else
  const char* self = "spherocyl_qscompress";
#endif
  try {
    int argn = 0;
    std::string file = string_input(argc, argv, ++argn, "Data file ('r' for random)");
    std::string dir = string_input(argc, argv, ++argn, "Output data directory");
    int n = int_input(argc, argv, ++argn, "Number of particles");
    double rate = float_input(argc, argv, ++argn, "Resize rate (epsilon)");
    double step = float_input(argc, argv, ++argn, "Integration step size");
#ifdef SPHEROCYL_FIND_JAM
    double target = float_input(argc, argv, ++argn, "Energy at jamming");
#else
    double target = float_input(argc, argv, ++argn, "Final packing fraction");
#endif
    double pos_rate = float_input(argc, argv, ++argn, "Position data save rate");
    double se_rate = float_input(argc, argv, ++argn, "Stress data save rate");
    (void)pos_rate;  // parsed and unused, as in the reference
    (void)se_rate;
    double pad = float_input(argc, argv, ++argn, "Cell padding");
    if (n <= 0) throw SpherocylError(SC_ERR_ARG, "number of particles must be positive");

    std::vector<double> x(n), y(n), phi(n), R(n), A(n);
    double L = 0, gamma = 0, gamma_total = 0;
    if (file == "r") {
      double packing = float_input(argc, argv, ++argn, "Packing Fraction");
      double aspect = float_input(argc, argv, ++argn, "Aspect ratio (i.e. A/R or 2A/D):");
      double bidisp = float_input(argc, argv, ++argn, "Bidispersity ratio: (1 for monodisperse):");
      int config = SC_CONFIG_RANDOM;
      if (argc > argn + 1)
        config = config_from_gpu_code(int_input(argc, argv, ++argn, "Initial configuration"));
      double area = 0;
      for (int p = 0; p < n; ++p) {
        R[p] = (p % 2 == 0) ? 0.5 : 0.5 * bidisp;
        A[p] = aspect * R[p];
        area += (4.0 * A[p] + kRefPi * R[p]) * R[p];
      }
      L = std::sqrt(area / packing);
      std::cout << "Box length L: " << L << std::endl;
      int st = sc_random_configuration(n, L, 0.0, A.data(), R.data(), config, seed_from_env(),
                                       100000, x.data(), y.data(), phi.data());
      if (st != SC_OK) throw SpherocylError(st, "random configuration failed");
    } else {
      PositionFile f = read_position_file(file);
      if (f.n != n) {
        std::cerr << "Warning: Number of particles in data file may not match requested number"
                  << std::endl;
        n = f.n;
      }
      x = f.x;
      y = f.y;
      phi = f.phi;
      R = f.R;
      A = f.A;
      L = f.L;
      gamma = f.gamma;
      gamma_total = f.gamma_total;
    }
    SpherocylBox box(L, n, x.data(), y.data(), phi.data(), A.data(), R.data(), pad);
    box.set_contact_model(SC_CONTACT_TWO_POINT);
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box.set_precision(true);
    box.set_gamma(gamma);
    box.set_total_gamma(gamma_total);
    box.set_step(step);
    box.set_output_directory(dir);
    box.set_se_file("sd_stress_energy.dat");  // the GPU class default (src/cuda/spherocyl_box.cpp:247)
    std::time_t start = std::time(nullptr);
    box.calc_se();
    box.display(0, 0);
#ifdef SPHEROCYL_FIND_JAM
    unsigned long t = box.find_jam_2p(target, rate);
#else
    unsigned long t = box.compress_qs2p(target, rate);
#endif
    long total = 0;
    for (int it : box.qs_iterations()) total += it;
    std::cout << "Steps: " << t << "  minimisations: " << box.qs_iterations().size()
              << "  FIRE iterations: " << total << "  packing: " << box.packing_fraction()
              << "  E: " << box.energy() << std::endl;
    box.display(0, 0);
    std::cout << "\nRun Time: " << std::time(nullptr) - start << std::endl;
  } catch (const SpherocylError& e) {
    std::cerr << self << ": " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
