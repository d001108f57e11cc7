// SpherocylBox::run_pure_strain and flip_shear_direction on a seeded random
// packing (compiled and run by tests/test_gpu_pure_shear.py).
//   box_pure_shear <output dir> <n> <packing> <seed> <rate> <dt> <stop aspect>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "spherocyl_box.h"

int main(int argc, char** argv) {
  if (argc < 8) return 2;
  const char* dir = argv[1];
  const int n = std::atoi(argv[2]);
  const double packing = std::atof(argv[3]);
  const unsigned long long seed = std::strtoull(argv[4], nullptr, 10);
  const double rate = std::atof(argv[5]), dt = std::atof(argv[6]), stop = std::atof(argv[7]);
  std::vector<double> x(n), y(n), phi(n), R(n), A(n);
  double area = 0;
  for (int p = 0; p < n; ++p) {
    R[p] = p % 2 == 0 ? 0.5 : 0.7;
    A[p] = R[p];
    area += (4.0 * A[p] + 3.14159265358979 * R[p]) * R[p];
  }
  const double L = std::sqrt(area / packing);
  if (sc_random_configuration(n, L, 0.0, A.data(), R.data(), SC_CONFIG_RANDOM, seed, 100000,
                              x.data(), y.data(), phi.data()) != SC_OK)
    return 3;
  try {
    SpherocylBox box(L, n, x.data(), y.data(), phi.data(), A.data(), R.data(), 0.1);
    box.set_contact_model(SC_CONTACT_TWO_POINT);
    box.set_output_directory(dir);
    box.save_positions_sd(999);  // the start, for the test's oracle run
    box.set_se_file("pure_se.dat");
    box.set_strain_rate(-rate);  // the sign is dropped, as in the reference
    box.set_step(dt);
    unsigned long t = box.run_pure_strain(0, stop, 10 * rate * dt, 0.25);
    box.flip_shear_direction();
    box.save_positions_sd(1000);
    std::printf("%lu %.17g\n", t, box.gamma());
  } catch (const SpherocylError& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
