// spherocyl_qscompress / spherocyl_findjam / spherocyl_ycompress /
// spherocyl_findjam_1p / spherocyl_qscycle: the quasi-static protocols with the
// argument order of the reference's GPU drivers.  qscompress, findjam,
// ycompress and findjam_1p follow src/cuda/main_2point_qscompress.cpp,
// main_2point_findjam.cpp, main_2point_ycompress.cpp and main_1point_findjam.cpp
// (lines 60-251):
//   1 data file ("r" for random)   2 output directory   3 N
//   4 resize rate (epsilon)   5 integration step size
//   6 final packing fraction (qscompress) | energy at jamming (findjam) |
//     target energy (ycompress)
//   7 position save rate   8 stress save rate   9 cell padding
//   [random only: 10 packing fraction  11 aspect ratio A/R  12 bidispersity
//    13 initial configuration (0 random, 2 random-aligned, 3 zero-energy,
//       5 zero-energy-aligned, 6 grid, 8 grid-aligned)]
// qscycle follows src/cuda/main_qs_compress.cpp (lines 89-268) and runs
// quasistatic_find_jam(max E, 1e-16, step):
//   1 data file   2 output directory   3 N   4 quasi-static resize step
//   5 minimum relaxation step (the reference's conjugate-gradient line-search
//     bound; parsed, unused by FIRE)   6 maximum energy during compression
//   7 number of cycles (parsed, unused, as in the reference)   8 stress save
//   rate   9 cell padding   [random only: 10-13 as above]
// Random starts: R = 0.5 and A = aspect R for even indices, both times the
// bidispersity for odd ones (src/cuda/main_2point_qscompress.cpp:164-170).
// A data file is either format: save_positions (sp*.dat, "N L ..." header)
// or the GPU code's sd*.dat ("N Lx Ly ..." header; Lx must equal Ly here).
// qscompress, findjam and ycompress use the two-point contact model, findjam_1p and
// qscycle the closest-point model (the reference's one-point protocols);
// minimisation is FIRE to F_rms < 1e-10.  SPHEROCYL_PRECISION=mixed selects
// FP32 pair forces; SPHEROCYL_SEED seeds random starts;
// SPHEROCYL_FIRE_MAX_ITER caps each minimisation (default 200000; a
// closest-point packing can stall on the alignment kink, where the
// reference's minimiser stops at its step limit too).
#include <cstdlib>
#include <ctime>
#include <iostream>
#include <vector>

#include "driver_common.h"
#include "spherocyl_files.h"


int main(int argc, char* argv[]) {
  using namespace driver;
#if defined(SPHEROCYL_QS_CYCLE)
  const char* self = "spherocyl_qscycle";
#elif defined(SPHEROCYL_FIND_JAM_1P)
  const char* self = "spherocyl_findjam_1p";
#elif defined(SPHEROCYL_Y_COMPRESS)
  const char* self = "spherocyl_ycompress";
#elif defined(SPHEROCYL_FIND_JAM)
  const char* self = "spherocyl_findjam";
#else
  const char* self = "spherocyl_qscompress";
#endif
  try {
    int argn = 0;
    std::string file = string_input(argc, argv, ++argn, "Data file ('r' for random)");
    std::string dir = string_input(argc, argv, ++argn, "Output data directory");
    int n = int_input(argc, argv, ++argn, "Number of particles");
#ifdef SPHEROCYL_QS_CYCLE
    double rate = float_input(argc, argv, ++argn, "Quasistatic resizing step size");
    double min_step = float_input(argc, argv, ++argn, "Minimum relaxation step size");
    double target = float_input(argc, argv, ++argn, "Maximum energy during compress");
    int cycles = int_input(argc, argv, ++argn, "Number of compression cycles");
    double se_rate = int_input(argc, argv, ++argn, "Stress data save rate");
    (void)min_step;  // parsed; FIRE has no line search
    (void)cycles;    // parsed and unused, as in the reference (main_qs_compress.cpp:268)
    (void)se_rate;
    const double step = 0.01;
#else
    double rate = float_input(argc, argv, ++argn, "Resize rate (epsilon)");
    double step = float_input(argc, argv, ++argn, "Integration step size");
#if defined(SPHEROCYL_FIND_JAM) || defined(SPHEROCYL_FIND_JAM_1P)
    double target = float_input(argc, argv, ++argn, "Energy at jamming");
#elif defined(SPHEROCYL_Y_COMPRESS)
    double target = float_input(argc, argv, ++argn, "Target Energy");
#else
    double target = float_input(argc, argv, ++argn, "Final packing fraction");
#endif
    double pos_rate = float_input(argc, argv, ++argn, "Position data save rate");
    double se_rate = float_input(argc, argv, ++argn, "Stress data save rate");
    (void)pos_rate;  // parsed and unused, as in the reference
    (void)se_rate;
#endif
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
#if defined(SPHEROCYL_QS_CYCLE) || defined(SPHEROCYL_FIND_JAM_1P)
    box.set_contact_model(SC_CONTACT_CLOSEST);
#else
    box.set_contact_model(SC_CONTACT_TWO_POINT);
#endif
    if (const char* mi = std::getenv("SPHEROCYL_FIRE_MAX_ITER")) {
      double fp[10] = {0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000};
      fp[9] = std::atof(mi);
      if (!(fp[9] >= 1)) throw SpherocylError(SC_ERR_ARG, "SPHEROCYL_FIRE_MAX_ITER must be >= 1");
      box.set_minimizer_params(fp);
    }
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box.set_precision(true);
    box.set_gamma(gamma);
    box.set_total_gamma(gamma_total);
    box.set_step(step);
    box.set_output_directory(dir);
    box.set_se_file("sd_stress_energy.dat");  // the GPU class default (src/cuda/spherocyl_box.cpp:247)
    std::time_t start = std::time(nullptr);
    box.calc_se();
    box.display(0, 0);
#if defined(SPHEROCYL_QS_CYCLE)
    unsigned long t = box.quasistatic_find_jam(target, 1e-16, rate);
#elif defined(SPHEROCYL_FIND_JAM_1P)
    unsigned long t = box.find_jam_1p(target, rate);
#elif defined(SPHEROCYL_FIND_JAM)
    unsigned long t = box.find_jam_2p(target, rate);
#elif defined(SPHEROCYL_Y_COMPRESS)
    unsigned long t = box.y_compress_energy_2p(target, rate);
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
