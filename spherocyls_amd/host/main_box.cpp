// spherocyl_box: the shear driver of the reference's GPU code, with its
// argument order (src/cuda/main.cpp:60-315, the order input/input_key.txt
// describes):
//   1 data file ("r" for random)   2 output directory   3 N   4 strain rate
//   5 integration step   6 run length (gamma)   7 position save rate
//   8 stress save rate   9 cell padding
//   10 change of shear direction: 0 none, 1 reverse (flip_shear_direction),
//      2 rotate by gamma (rotate_by_gamma); 3 (shear past the remap to a given
//      strain, then rotate) is not built and stops with an error
//   [random only: 11 packing fraction  12 aspect ratio A/R  13 bidispersity
//    [14 initial configuration (0 random, 2 random aligned, 3 zero energy,
//        5 zero energy aligned, 6 grid, 8 grid aligned)
//     15 relative width of the aspect-ratio distribution (0: every rod the
//        same aspect; > 0: gamma-distributed aspect ratios)]]
// Random starts: R = 0.5 for even and 0.5 bidispersity for odd indices,
// A = aspect R (src/cuda/main.cpp:157-166); with a distribution width the
// aspect ratios come from sc_gamma_shapes and the box is sized for the drawn
// shapes.  A data file is an sd*.dat (square boxes) or sp*.dat file; the run
// restarts at the 10-digit time stamp of its name and seeks the SE file.
// Output: the GPU class's files, sd_stress_energy.dat lines
// "t E Pxx Pyy P Pxy Pyx" and sd%010lu.dat positions (run_strain_sd).
// SPHEROCYL_SEED seeds random starts; SPHEROCYL_PRECISION=mixed selects FP32
// pair forces; SPHEROCYL_CONTACT=two_point the two-point contact model.
//
// Built with SPHEROCYL_PURE_SHEAR it is spherocyl_pureshear, the pure-shear
// driver (src/cuda/main_pure_shear.cpp:60-296): argument 6 is the box aspect
// Lx / Ly at which pure shear stops, argument 10 a rotation by gamma (1)
// before the run, and the run is run_pure_strain with the two-point contact
// model, as run_pure_strain_2p there.
#include <cmath>
#include <ctime>
#include <iostream>
#include <vector>

#include "driver_common.h"
#include "spherocyl_files.h"

static const char* self_name() {
#ifdef SPHEROCYL_PURE_SHEAR
  return "spherocyl_pureshear";
#else
  return "spherocyl_box";
#endif
}

int main(int argc, char* argv[]) {
  using namespace driver;
  try {
    int argn = 0;
    std::string file = string_input(argc, argv, ++argn, "Data file ('r' for random)");
    std::string dir = string_input(argc, argv, ++argn, "Output data directory");
    int n = int_input(argc, argv, ++argn, "Number of particles");
    double rate = float_input(argc, argv, ++argn, "Strain rate");
    double step = float_input(argc, argv, ++argn, "Integration step size");
#ifdef SPHEROCYL_PURE_SHEAR
    double run = float_input(argc, argv, ++argn, "Strain stop (Box Aspect Lx/Ly)");
#else
    double run = float_input(argc, argv, ++argn, "Strain run length (gamma)");
#endif
    double pos_rate = float_input(argc, argv, ++argn, "Position data save rate");
    double se_rate = float_input(argc, argv, ++argn, "Stress data save rate");
    double pad = float_input(argc, argv, ++argn, "Cell padding");
#ifdef SPHEROCYL_PURE_SHEAR
    int change = int_input(argc, argv, ++argn, "Change shear direction (1 to reverse");
    if (change == 1) change = 2;  // main_pure_shear.cpp:261-265 rotates by gamma
    else if (change != 0) throw SpherocylError(SC_ERR_ARG, "change of shear direction must be 0 or 1");
#else
    int change = int_input(argc, argv, ++argn,
                           "Change shear direction (1 to reverse, 2 to rotate by gamma, 3 to "
                           "increase gama and rotate)");
    if (change == 3)
      throw SpherocylError(SC_ERR_ARG,
                           "shearing past the Lees-Edwards remap before a rotation (3) is not "
                           "supported; 0, 1 and 2 are");
    if (change < 0 || change > 3)
      throw SpherocylError(SC_ERR_ARG, "change of shear direction must be 0, 1 or 2");
#endif
    if (n <= 0) throw SpherocylError(SC_ERR_ARG, "number of particles must be positive");
    if (se_rate < rate * step) se_rate = rate * step;
    if (pos_rate < rate) pos_rate = rate;

    std::vector<double> x(n), y(n), phi(n), R(n), A(n);
    double L = 0, gamma = 0, gamma_total = 0;
    unsigned long t0 = 0;
    if (file == "r") {
      double packing = float_input(argc, argv, ++argn, "Packing Fraction");
      double aspect = float_input(argc, argv, ++argn, "Aspect ratio (i.e. A/R or 2A/D)");
      double bidisp = float_input(argc, argv, ++argn, "Bidispersity ratio: (1 for monodisperse)");
      int config = SC_CONFIG_RANDOM;
      double sigma = 0.0;
      if (argc > argn + 1) {
        config = config_from_gpu_code(int_input(argc, argv, ++argn, "Initial configuration"));
        sigma = float_input(argc, argv, ++argn,
                            "Aspect Ratio distribution width (relative deviation)");
      }
      if (sigma > 0) {
        int st = sc_gamma_shapes(n, aspect, sigma, bidisp, seed_from_env(), R.data(), A.data());
        if (st != SC_OK) throw SpherocylError(st, "gamma-distributed shapes failed");
      } else {
        for (int p = 0; p < n; ++p) {
          R[p] = (p % 2 == 0) ? 0.5 : 0.5 * bidisp;
          A[p] = aspect * R[p];
        }
      }
      double area = 0;
      for (int p = 0; p < n; ++p) area += (4.0 * A[p] + kRefPi * R[p]) * R[p];
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
      if (f.strain_rate != rate)
        std::cerr << "Warning: Strain rate in data file does not match the requested rate"
                  << std::endl;
      if (f.step != step)
        std::cerr << "Warning: Integration step size in data file does not match requested step"
                  << std::endl;
      x = f.x;
      y = f.y;
      phi = f.phi;
      R = f.R;
      A = f.A;
      L = f.L;
      gamma = f.gamma;
      gamma_total = f.gamma_total;
      t0 = position_file_time(file);
    }
    SpherocylBox box(L, n, x.data(), y.data(), phi.data(), A.data(), R.data(), pad);
#ifdef SPHEROCYL_PURE_SHEAR
    box.set_contact_model(SC_CONTACT_TWO_POINT);
#else
    if (env_flag("SPHEROCYL_CONTACT", "two_point")) box.set_contact_model(SC_CONTACT_TWO_POINT);
#endif
    if (env_flag("SPHEROCYL_PRECISION", "mixed")) box.set_precision(true);
    box.set_gamma(gamma);
    box.set_total_gamma(gamma_total);
    box.set_strain_rate(rate);
    box.set_step(step);
    box.set_output_directory(dir);
    box.set_se_file("sd_stress_energy.dat");  // the GPU class default (src/cuda/spherocyl_box.cpp:247)
    if (change == 1) {
      box.set_total_gamma(0);
      box.flip_shear_direction();
      std::cout << "Shear direction reversed" << std::endl;
    } else if (change == 2) {
      box.set_total_gamma(0);
      box.rotate_by_gamma();
      std::cout << "Shear direction rotated" << std::endl;
    }
    std::time_t start = std::time(nullptr);
    box.calc_se();
    box.display(0, 0);
#ifdef SPHEROCYL_PURE_SHEAR
    unsigned long t = box.run_pure_strain(t0, run, se_rate, pos_rate);
#else
    unsigned long t = box.run_strain_sd(t0, run, se_rate, pos_rate);
#endif
    box.display(0, 0);
    std::cout << "Steps: " << t << "\nRun Time: " << std::time(nullptr) - start << std::endl;
  } catch (const SpherocylError& e) {
    std::cerr << self_name() << ": " << e.what() << std::endl;
    return 1;
  }
  return 0;
}
