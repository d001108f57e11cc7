// spherocyl_box.h -- the reference's simulation-box class, MI355X edition.
//
// SpherocylBox keeps the public surface of the reference's CPU class
// (src/cpp/spherocyl_box.h:85-120): the same constructors, the same method
// names and argument meanings, and the same output files (save_positions,
// the *_se.dat stress/energy files of run_strain, resize_box and
// gradient_descent_minimize, byte-for-byte in layout and number formatting).
// Every force pass and integration step runs on the GPU through the C ABI of
// spherocyls.h; the class holds a host mirror of the particle arrays that is
// refreshed only when a caller or a file writer needs it.
//
// Differences from the reference, all deliberate:
//   - errors throw SpherocylError instead of assert()/exit(1);
//   - random_configuration is seeded (set_seed), the reference uses time(0);
//   - read accessors for forces, torques, energy, stress, contacts and
//     neighbours, which the reference lacks (it only prints or writes files);
//   - FIRE minimisation (fire_minimize), FP32-pair precision (set_precision);
//   - resize_box follows the reference's predictor, which takes the minimum
//     image with the box length of the previous sub-step
//     (src/cpp/resize_dynamics.cpp:17-86, 101); set_resize_image(true) uses the
//     current length instead.
#ifndef SPHEROCYLS_SPHEROCYL_BOX_H
#define SPHEROCYLS_SPHEROCYL_BOX_H

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "spherocyls.h"

// src/cpp/spherocyl_box.h:8 plus the two grid modes of spherocyl_box.cpp:718-737
enum initialConfig { RANDOM, RANDOM_ALIGNED, ZERO_E, ZERO_E_ALIGNED, GRID, GRID_ALIGNED };

class SpherocylError : public std::runtime_error {
 public:
  SpherocylError(int status, const std::string& what) : std::runtime_error(what), m_nStatus(status) {}
  int status() const { return m_nStatus; }

 private:
  int m_nStatus;
};

class SpherocylBox {
 public:
  // src/cpp/spherocyl_box.cpp:57-173.  The array constructors copy the arrays.
  SpherocylBox();
  SpherocylBox(double dL, int nParticles);
  SpherocylBox(double dL, int nParticles, double dPadding);
  SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi, double dA,
               double dR);
  SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi, double dA,
               double dR, double dPadding);
  SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi, double* pdA,
               double* pdR);
  SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi, double* pdA,
               double* pdR, double dPadding);
  ~SpherocylBox();
  SpherocylBox(const SpherocylBox&) = delete;
  SpherocylBox& operator=(const SpherocylBox&) = delete;

  // src/cpp/spherocyl_box.cpp:188-221: wraps x, y into [0, L) and phi into
  // [-pi, pi) as the reference does.
  void set_particle(int nIndex, double dX, double dY, double dPhi, double dA, double dR);
  // Cell grid and neighbour bookkeeping live on the GPU and are exact on every
  // pass; these calls only make the device state current
  // (src/cpp/spherocyl_box.cpp:223-355).
  void configure_cells();
  void find_cells();
  void find_neighbors();
  // src/cpp/spherocyl_box.cpp:368-408 (stdout)
  void display(bool bParticles = 0, bool bCells = 0);
  // src/cpp/spherocyl_box.cpp:410-423: <output dir>/<strFile>
  void save_positions(std::string strFile);
  // src/cpp/spherocyl_box.cpp:502-522, strain_dynamics.cpp:137-156
  void calc_se();
  void calc_forces();
  // src/cpp/spherocyl_box.cpp:694-790 (seeded; see set_seed)
  void random_configuration(double dA, double dR, initialConfig config = RANDOM);
  // src/cpp/strain_dynamics.cpp:238-393
  long unsigned int run_strain(double dRunLength);
  long unsigned int run_strain(double dRunLength, long unsigned int nTime);
  // src/cpp/resize_dynamics.cpp:244-416
  long unsigned int resize_box(double dFinalPacking);
  long unsigned int resize_box(double dFinalPacking, long unsigned int nTime);
  // src/cpp/relaxation_dynamics.cpp:33-85
  void gradient_descent_minimize(int nMaxSteps = 1000000, double dMinE = 0);

  // src/cpp/spherocyl_box.h:110-119
  void set_output_directory(std::string strOutputDir) { m_strOutputDir = strOutputDir; }
  void set_se_file(std::string strSEOutput) { m_strSEOutput = strSEOutput; }
  void set_strain_rate(double dStrainRate) { m_dStrainRate = dStrainRate; }
  void set_resize_rate(double dResizeRate) { m_dResizeRate = dResizeRate; }
  void set_step(double dStep) { m_dStep = dStep; }
  void set_pos_save_step(double dPosSaveStep) { m_dPosSaveStep = dPosSaveStep; }
  void set_se_save_step(double dSESaveStep) { m_dSESaveStep = dSESaveStep; }
  void set_cell_padding(double dPadding);
  void set_gamma(double dGamma);
  void set_total_gamma(double dTotalGamma);

  // ---- additions (no reference counterpart) ----
  // FIRE to F_rms < ftol.  params: SC ABI order (dt0, dt_max, n_min, f_inc,
  // f_dec, alpha0, f_alpha, max_move, ftol, max_iter); nullptr = defaults.
  // Returns the iterations taken; throws on SC_ERR_NOT_CONVERGED only if
  // bThrowIfNotConverged.
  int fire_minimize(const double* pdParams = nullptr, bool bThrowIfNotConverged = false);
  void set_precision(bool bMixed);
  // SC_CONTACT_CLOSEST (the CPU reference's model, the default) or
  // SC_CONTACT_TWO_POINT (src/cuda/2point_dynamics.cu:391-546).
  void set_contact_model(int nModel);
  // FIRE parameters (SC ABI order) of the quasi-static protocols below;
  // nullptr restores the defaults.
  void set_minimizer_params(const double* pdParams);

  // Quasi-static protocols of the reference's GPU code
  // (src/cuda/2point_dynamics.cu:4300-4632).  A step multiplies x, y and L by
  // 1 - eps (compress) or 1 / (1 - eps) (expand), as resize_step_2p does
  // (src/cuda/2point_dynamics.cu:1187-1285), takes one unsheared Heun step of
  // dt 0.01 (the box's step is 0.01 afterwards, as in the reference), then
  // minimises to F_rms < ftol with FIRE (the reference runs its
  // conjugate-gradient minimiser).  After
  // each step: a line "t packing E Pxx Pyy Pxy Pyx" in the SE file (the
  // reference's printf formats; the stress is the corrected virial with the
  // CPU code's sign) and a position file sd%010lu.dat in the GPU code's text
  // format, every step by default (set_qs_position_interval).  Each protocol
  // runs with the box's contact model and returns the last step number.
  //   compress_qs2p: minimise, then compress while packing < dMaxPacking.
  //   expand_qs2p:   minimise, then expand while packing > dMinPacking.
  //   find_jam_2p:   bracket the packing where the minimised energy crosses
  //                  dJamE: compress (E < dJamE) or expand (E > dJamE) with
  //                  steps growing by 1.1 until E crosses, then continue with
  //                  step * 1.1^(steps - 1) / 2, until the step is below 1e-15;
  //                  the final state also goes to sd%010lu.bin.
  //   y_compress_energy_2p: the search of find_jam_2p towards dTargetE with
  //                  y-only steps (y and Ly scaled, x and Lx kept) down to a
  //                  step of 1e-14.
  // run_pure_strain_2p (src/cuda/2point_dynamics.cu:2286-2397): pure shear
  // (x expands, y contracts at half the strain rate; sc_pure_strain_steps) at
  // |strain rate| until Lx / Ly reaches dStopAspect.  Every dSvStressGamma of
  // strain a line "t E Pxx Pyy P Pxy Pyx" (printf %.7g; corrected stress with
  // the CPU code's sign, P = (Pxx + Pyy) / 2) of the state the step starts
  // from; every dSvPosGamma a position file sd%010lu.dat.  Restarts at nStart
  // seek the SE file as run_strain does.  The reference ends with one more
  // step of its Lees-Edwards integrator before the final save; here the final
  // state is saved as it is.  Returns the last time stamp.
  long unsigned int run_pure_strain(long unsigned int nStart, double dStopAspect,
                                    double dSvStressGamma, double dSvPosGamma);
  // run_strain of the reference's GPU class (src/cuda/strain_dynamics.cu:
  // 984-1083): sheared Heun steps from time nStart for dRunGamma of strain,
  // with the GPU class's files: SE lines and sd%010lu.dat positions as for
  // run_pure_strain above.  Returns the last time stamp.
  long unsigned int run_strain_sd(long unsigned int nStart, double dRunGamma,
                                  double dSvStressGamma, double dSvPosGamma);
  // flip_shear_direction (src/cuda/find_neighbors.cu:565-579): x -> L - x,
  // phi -> -phi, gamma -> -gamma
  void flip_shear_direction();
  // rotate_by_gamma (src/cuda/find_neighbors.cu:624-645): rotate the frame so
  // the sheared box vector is the new x axis; the box becomes non-square
  void rotate_by_gamma();
  // One-point quasi-static protocols of the reference's GPU code
  // (src/cuda/energy_minimize.cu:1731-1961).  Steps as above, except that an
  // expansion multiplies by 1 + eps (resize_step, src/cuda/strain_dynamics.cu:
  // 1124-1196).  The reference runs them with the
  // closest-point contact model and its conjugate-gradient minimiser with
  // angle damping; here each minimisation is FIRE with the box's contact model
  // and minimiser parameters (set_minimizer_params: a closest-point packing
  // that stalls on the alignment kink stops at max_iter, as the reference's
  // minimiser stops at its step limit).  SE lines "t packing E Pxx Pyy Pxy Pyx"
  // and sd%010lu.dat position files as for the two-point protocols.
  //   quasistatic_compress: minimise; expand while E > 0; compress while
  //                         packing < dMaxPack.
  //   qs_one_cycle:         minimise; expand while E > 0; compress while
  //                         E < dMaxE (SE lines to the open file f); returns
  //                         the last time.
  //   quasistatic_cycle:    nCycles cycles.
  //   quasistatic_find_jam: cycles with dMaxE and the step halved after each,
  //                         while dMaxE > dMinE and the step > Lx 1e-15.
  //   find_jam_1p:          steps growing by 1.1 towards the energy dJamE,
  //                         the next search from step 1.1^(steps) / 2, until
  //                         the step is below Lx 1e-13; final state also to
  //                         sd%010lu.bin.
  void quasistatic_compress(double dMaxPack, double dResizeStep);
  long unsigned int qs_one_cycle(FILE* f, long unsigned int nTime, double dMaxE,
                                 double dResizeStep);
  long unsigned int quasistatic_cycle(int nCycles, double dMaxE, double dResizeStep);
  long unsigned int quasistatic_find_jam(double dMaxE, double dMinE, double dResizeStep);
  long unsigned int find_jam_1p(double dJamE, double dResizeStep);
  long unsigned int compress_qs2p(double dMaxPacking, double dResizeStep);
  long unsigned int expand_qs2p(double dMinPacking, double dResizeStep);
  long unsigned int find_jam_2p(double dJamE, double dResizeStep);
  long unsigned int y_compress_energy_2p(double dTargetE, double dResizeStep);
  void set_qs_position_interval(int nSteps) { m_nQsPosEvery = nSteps; }
  // iterations of every minimisation of the last quasi-static protocol
  const std::vector<int>& qs_iterations() const { return m_vQsIters; }
  // save_positions_bin / the GPU code's save_positions (sd%010lu.bin / .dat)
  void save_positions_bin(unsigned long nTime);
  void save_positions_sd(unsigned long nTime);
  void set_seed(uint64_t nSeed) { m_nSeed = nSeed; }
  void set_device(int nDevice) { m_nDevice = nDevice; }
  void set_resize_image(bool bCurrentLength) { m_bResizeCorrected = bCurrentLength; }

  int particles() const { return m_nParticles; }
  double box_length() const { return m_dLx; }
  double packing_fraction() const { return m_dPacking; }
  double gamma();
  double total_gamma();
  double energy() const { return m_dEnergy; }
  // reference formulas (src/cpp/spherocyl_box.cpp:483-495): Pxx, Pyy, Pxy
  void stress(double* pdPxx, double* pdPyy, double* pdPxy) const;
  // corrected virial: Pxx, Pyy, Pxy, Pyx
  void corrected_stress(double* pdP) const;
  double residual_force();  // F_rms of the last force pass
  int contact_count() const { return m_nContacts; }
  void positions(std::vector<double>& x, std::vector<double>& y, std::vector<double>& phi);
  void forces(std::vector<double>& fx, std::vector<double>& fy, std::vector<double>& tau);
  std::vector<std::pair<int, int>> contacts();
  std::vector<std::pair<int, int>> neighbors();
  sc_system* handle();  // the underlying C ABI handle (device state current)

 private:
  void init(double dL, int nParticles, double dPadding);
  void check(int status, const char* where);
  void ensure_device();
  void sync_host();
  void read_scalars();
  void write_se_line(std::ostream& out, unsigned long t, bool bPacking);
  double packing_from_arrays() const;
  int qs_minimize();
  void qs_step(double dEpsilon, bool bTwoPoint, bool bYOnly = false);
  void qs_line(FILE* f, const char* fmt, unsigned long t);
  long unsigned int qs_run(double dLimit, double dResizeStep, bool bCompress);
  FILE* qs_open_se();
  long unsigned int resize_to_energy_1p(FILE* f, long unsigned int nTime, double dEnergy,
                                        double dStep);
  long unsigned int resize_to_energy_2p(FILE* f, long unsigned int nTime, double dEnergy,
                                        double dStep, bool bYOnly);
  long unsigned int run_sd(long unsigned int nStart, bool bPure, unsigned long nStop,
                           double dStopAspect, double dSvStressGamma, double dSvPosGamma,
                           const char* name);
  long unsigned int jam_search_2p(double dJamE, double dResizeStep, double dMinResizeStep,
                                  bool bYOnly, const char* name);

  int m_nParticles = 0;
  double m_dLx = 0, m_dLy = 0;
  double m_dGamma = 0, m_dGammaTotal = 0;
  double m_dStrainRate = 0, m_dResizeRate = 0, m_dStep = 0.05;
  double m_dPosSaveStep = 0.01, m_dSESaveStep = 1e-5;
  double m_dPacking = 0, m_dPadding = 0;
  double m_dEnergy = 0, m_dPxx = 0, m_dPyy = 0, m_dPxy = 0;
  double m_dCorr[4] = {0, 0, 0, 0};
  double m_dFrms = 0;
  int m_nContacts = 0;
  std::vector<double> m_vX, m_vY, m_vPhi, m_vR, m_vA;
  std::string m_strOutputDir = ".";
  std::string m_strSEOutput = "sp_se.dat";

  sc_system* m_sys = nullptr;
  int m_nDevice = 0;
  int m_nPrecision = SC_PREC_FP64;
  uint64_t m_nSeed = 1;
  bool m_bResizeCorrected = false;
  int m_nContactModel = SC_CONTACT_CLOSEST;
  double m_dFire[10] = {0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000};
  int m_nQsPosEvery = 1;
  std::vector<int> m_vQsIters;
  bool m_bGeomStale = true;   // radii, lengths, box or padding changed: rebuild the system
  bool m_bDevStale = true;    // host positions newer than the device
  bool m_bHostStale = false;  // device positions newer than the host mirror
};

// Position files (save_positions format, src/cpp/spherocyl_box.cpp:410-423):
// one header line "N L packing gamma gamma_total strain_rate step", then
// "x y phi R A" per particle, 14 significant digits.
struct PositionFile {
  int n = 0;
  double L = 0, packing = 0, gamma = 0, gamma_total = 0, strain_rate = 0, step = 0;
  std::vector<double> x, y, phi, R, A;
};
// Reads a save_positions file (the restart format read by src/cpp/main.cpp:130-157
// through DatFileInput, src/cpp/dat_file_input.cpp:36-58).  Throws on error.
PositionFile read_position_file(const std::string& path);
// The time stamp of an sp*/sr* file name: the 10 digits before ".dat"
// (src/cpp/main.cpp:152-155).
unsigned long position_file_time(const std::string& path);

#endif
