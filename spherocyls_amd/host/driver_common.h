// Shared argument handling of the command-line drivers.
//
// Arguments come from argv, or, when argv runs out, one per line from stdin
// (so "driver < input_file.txt" works), each echoed with its description, as
// the reference's int_input / float_input / string_input do
// (src/cpp/main.cpp:20-59).
//
// Environment (no reference counterpart):
//   SPHEROCYL_SEED=<n>          seed of random starting configurations (default 1)
//   SPHEROCYL_PRECISION=mixed   FP32 pair math with FP64 sums
//   SPHEROCYL_RESIZE_IMAGE=current  resize predictor with the current box length
//   SPHEROCYL_APPLY_SAVE_RATES=1  pass the position / stress save rates to the
//                               box (the reference parses and clamps them but
//                               never passes them on, src/cpp/main.cpp:87-90,
//                               181-184, so by default the box defaults apply)
#ifndef SPHEROCYLS_DRIVER_COMMON_H
#define SPHEROCYLS_DRIVER_COMMON_H

#include <cmath>
#include <cstdlib>
#include <iostream>
#include <string>

#include "spherocyl_box.h"

namespace driver {

// the reference's truncated pi (src/cpp/main.cpp:18)
constexpr double kRefPi = 3.14159265358979;

inline std::string string_input(int argc, char* argv[], int argn, const char* what) {
  std::string v;
  if (argc > argn) {
    v = argv[argn];
    std::cout << what << ": " << v << std::endl;
  } else {
    std::cout << what << ": ";
    std::cin >> v;
  }
  return v;
}

inline int int_input(int argc, char* argv[], int argn, const char* what) {
  if (argc > argn) {
    int v = std::atoi(argv[argn]);
    std::cout << what << ": " << v << std::endl;
    return v;
  }
  int v = 0;
  std::cout << what << ": ";
  std::cin >> v;
  return v;
}

inline double float_input(int argc, char* argv[], int argn, const char* what) {
  if (argc > argn) {
    double v = std::atof(argv[argn]);
    std::cout << what << ": " << v << std::endl;
    return v;
  }
  double v = 0;
  std::cout << what << ": ";
  std::cin >> v;
  return v;
}

inline unsigned long long seed_from_env() {
  const char* s = std::getenv("SPHEROCYL_SEED");
  return s ? std::strtoull(s, nullptr, 10) : 1ull;
}

inline bool env_flag(const char* name, const char* value) {
  const char* s = std::getenv(name);
  return s && std::string(s) == value;
}

// Box side for N particles of half-shaft A and radius R at packing fraction phi
// (src/cpp/main.cpp:113-114).
inline double box_side(int n, double A, double R, double phi) {
  double area = n * (4.0 * A + kRefPi * R) * R;
  return std::sqrt(area / phi);
}

// The GPU drivers' initial-configuration codes (src/cuda/main.cpp:120) mapped
// to sc_random_configuration's; the "uniform" angle variants (1, 4, 7) and the
// custom configuration (9) are not built.
inline int config_from_gpu_code(int c) {
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

}  // namespace driver

#endif
