// SpherocylBox::quasistatic_find_jam and quasistatic_compress (the reference's
// one-point quasi-static protocols) on a seeded random packing, closest-point
// contacts (compiled and run by tests/test_gpu_quasistatic.py).
//   box_qs_1p <output dir> <n> <packing> <seed> <step> <max E> <min E> <max pack>
//             <output dir of the compression>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "spherocyl_box.h"

int main(int argc, char** argv) {
  if (argc < 10) return 2;
  const char* dir = argv[1];
  const int n = std::atoi(argv[2]);
  const double packing = std::atof(argv[3]);
  const unsigned long long seed = std::strtoull(argv[4], nullptr, 10);
  const double step = std::atof(argv[5]), emax = std::atof(argv[6]), emin = std::atof(argv[7]);
  const double maxpack = std::atof(argv[8]);
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
    double fp[10] = {0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 2000};
    SpherocylBox box(L, n, x.data(), y.data(), phi.data(), A.data(), R.data(), 0.1);
    box.set_contact_model(SC_CONTACT_CLOSEST);
    box.set_minimizer_params(fp);
    box.set_output_directory(dir);
    box.set_qs_position_interval(5);
    box.set_se_file("cycle_se.dat");
    unsigned long t1 = box.quasistatic_find_jam(emax, emin, step);
    const double phi1 = box.packing_fraction();
    box.set_output_directory(argv[9]);
    box.set_se_file("compress_se.dat");
    box.quasistatic_compress(maxpack, step);
    std::printf("%lu %.17g %.17g\n", t1, phi1, box.packing_fraction());
  } catch (const SpherocylError& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
