// SpherocylBox on the GPU: the reference's box protocol (file output, cadences,
// restart) on the host, every force pass and step through the C ABI.
//
// Protocol loops follow the reference's, including its output quirks:
//   - run_strain writes an SE line after every nSESaveT-th sub-step of an outer
//     step and again at the end of every outer step t with t*nIntSteps a
//     multiple of nSESaveT, repeating the last computed values
//     (src/cpp/strain_dynamics.cpp:285-299); resize_box likewise
//     (src/cpp/resize_dynamics.cpp:293-306, 378-389);
//   - sub-step counters restart at 1 in every outer step, so an SE cadence that
//     does not divide the sub-steps per unit time restarts each outer step;
//   - the SE-save and position-save cadences are rounded to whole sub-steps /
//     outer steps with int(x + 0.5).
// Cadences that round to zero make the reference divide by zero; here they
// throw.
#include "spherocyl_box.h"
#include "spherocyl_files.h"

#include <cstring>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>

namespace {

// the reference's truncated pi (src/cpp/spherocyl_box.cpp:12)
constexpr double kRefPi = 3.14159265358979;
// SE records held per batched GPU call
constexpr long kMaxRecordsPerCall = 65536;

std::string time_name(const char* prefix, unsigned long t) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%s%010lu.dat", prefix, t);
  return buf;
}

// Positions the stream's write pointer just after the line whose leading
// integer equals t (restart of run_strain, src/cpp/strain_dynamics.cpp:250-279).
// Returns false when no such line exists.
bool seek_after_time(std::fstream& f, unsigned long t) {
  std::string line;
  while (std::getline(f, line)) {
    std::size_t sp = line.find(' ');
    if (std::strtoul(line.substr(0, sp).c_str(), nullptr, 10) == t && !line.empty()) {
      std::streampos pos = f.tellg();
      f.clear();
      if (pos == std::streampos(-1)) {  // last line had no newline: append one
        f.seekp(0, std::ios::end);
        f << "\n";
      } else {
        f.seekp(pos);
      }
      return true;
    }
  }
  return false;
}

}  // namespace

// "t v0 v1 ...\n" with the stream's own precision
static void se_line(std::ostream& out, unsigned long t, int nvals, const double* vals);

// ---------------------------------------------------------------- construction

SpherocylBox::SpherocylBox() {}

SpherocylBox::SpherocylBox(double dL, int nParticles) { init(dL, nParticles, 0.0); }

SpherocylBox::SpherocylBox(double dL, int nParticles, double dPadding) {
  init(dL, nParticles, dPadding);
}

SpherocylBox::SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi,
                           double dA, double dR)
    : SpherocylBox(dL, nParticles, pdX, pdY, pdPhi, dA, dR, 0.0) {}

SpherocylBox::SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi,
                           double dA, double dR, double dPadding) {
  init(dL, nParticles, dPadding);
  for (int p = 0; p < nParticles; ++p) set_particle(p, pdX[p], pdY[p], pdPhi[p], dA, dR);
  m_dPacking = packing_from_arrays();
}

SpherocylBox::SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi,
                           double* pdA, double* pdR)
    : SpherocylBox(dL, nParticles, pdX, pdY, pdPhi, pdA, pdR, 0.0) {}

SpherocylBox::SpherocylBox(double dL, int nParticles, double* pdX, double* pdY, double* pdPhi,
                           double* pdA, double* pdR, double dPadding) {
  init(dL, nParticles, dPadding);
  for (int p = 0; p < nParticles; ++p) set_particle(p, pdX[p], pdY[p], pdPhi[p], pdA[p], pdR[p]);
  m_dPacking = packing_from_arrays();
}

SpherocylBox::~SpherocylBox() {
  if (m_sys) sc_system_destroy(m_sys);
}

void SpherocylBox::init(double dL, int nParticles, double dPadding) {
  if (!(dL > 0) || nParticles < 0 || !(dPadding >= 0))
    throw SpherocylError(SC_ERR_ARG, "SpherocylBox: need L > 0, N >= 0, padding >= 0");
  m_nParticles = nParticles;
  m_dLx = m_dLy = dL;
  m_dPadding = dPadding;
  m_vX.assign(nParticles, 0.0);
  m_vY.assign(nParticles, 0.0);
  m_vPhi.assign(nParticles, 0.0);
  m_vR.assign(nParticles, 0.5);
  m_vA.assign(nParticles, 0.0);
  m_bGeomStale = true;
}

void SpherocylBox::check(int status, const char* where) {
  if (status == SC_OK) return;
  std::string msg = std::string(where) + ": " + sc_status_string(status);
  if (m_sys) {
    const char* e = sc_last_error(m_sys);
    if (e && *e) msg += std::string(" (") + e + ")";
  }
  throw SpherocylError(status, msg);
}

double SpherocylBox::packing_from_arrays() const {
  // calculate_packing_fraction, src/cpp/spherocyl_box.cpp:357-366
  double filled = 0.0;
  for (int p = 0; p < m_nParticles; ++p) filled += (kRefPi * m_vR[p] + 4.0 * m_vA[p]) * m_vR[p];
  return filled / (m_dLx * m_dLy);
}

void SpherocylBox::set_particle(int p, double dX, double dY, double dPhi, double dA, double dR) {
  if (p < 0 || p >= m_nParticles || !(dA >= 0.0) || !(dR > 0.0))
    throw SpherocylError(SC_ERR_ARG, "set_particle: index out of range or A < 0 or R <= 0");
  sync_host();
  // wrapping as src/cpp/spherocyl_box.cpp:204-215
  while (dX > m_dLx) dX -= m_dLx;
  while (dX < 0.0) dX += m_dLx;
  while (dY > m_dLy) dY -= m_dLy;
  while (dY < 0.0) dY += m_dLy;
  while (dPhi > kRefPi) dPhi -= 2 * kRefPi;
  while (dPhi < -kRefPi) dPhi += 2 * kRefPi;
  m_vX[p] = dX;
  m_vY[p] = dY;
  m_vPhi[p] = dPhi;
  if (m_vA[p] != dA || m_vR[p] != dR) m_bGeomStale = true;
  m_vA[p] = dA;
  m_vR[p] = dR;
  m_bDevStale = true;
}

void SpherocylBox::set_cell_padding(double dPadding) {
  if (!(dPadding >= 0)) throw SpherocylError(SC_ERR_ARG, "set_cell_padding: padding < 0");
  sync_host();
  m_dPadding = dPadding;
  m_bGeomStale = true;
}

void SpherocylBox::set_gamma(double dGamma) {
  sync_host();
  m_dGamma = dGamma;
  m_bDevStale = true;
}

void SpherocylBox::set_total_gamma(double dTotalGamma) {
  sync_host();
  m_dGammaTotal = dTotalGamma;
  m_bDevStale = true;
}

void SpherocylBox::set_precision(bool bMixed) {
  m_nPrecision = bMixed ? SC_PREC_MIXED : SC_PREC_FP64;
  if (m_sys) check(sc_set_precision(m_sys, m_nPrecision), "set_precision");
}

void SpherocylBox::set_contact_model(int nModel) {
  if (nModel != SC_CONTACT_CLOSEST && nModel != SC_CONTACT_TWO_POINT)
    throw SpherocylError(SC_ERR_ARG, "set_contact_model: unknown model");
  m_nContactModel = nModel;
  if (m_sys) check(sc_set_contact_model(m_sys, m_nContactModel), "set_contact_model");
}

void SpherocylBox::set_minimizer_params(const double* pdParams) {
  static const double kDefaults[10] = {0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000};
  std::copy(pdParams ? pdParams : kDefaults, (pdParams ? pdParams : kDefaults) + 10, m_dFire);
}

// ------------------------------------------------------------ device state

void SpherocylBox::ensure_device() {
  if (m_nParticles <= 0) throw SpherocylError(SC_ERR_STATE, "box has no particles");
  if (!m_sys || m_bGeomStale) {
    sync_host();
    if (m_dLx != m_dLy)
      throw SpherocylError(SC_ERR_STATE,
                           "SpherocylBox: a sheared-out (non-square) box cannot be rebuilt");
    if (m_sys) {
      sc_system_destroy(m_sys);
      m_sys = nullptr;
    }
    int offs[2] = {0, m_nParticles};
    double L = m_dLx, pad = m_dPadding;
    sc_system* s = nullptr;
    int st = sc_system_create(&s, m_nDevice, 1, offs, &L, m_vX.data(), m_vY.data(), m_vPhi.data(),
                              m_vA.data(), m_vR.data(), &pad);
    if (st != SC_OK) {
      std::string msg = std::string("SpherocylBox: ") + sc_status_string(st);
      if (s) {
        msg += std::string(" (") + sc_last_error(s) + ")";
        sc_system_destroy(s);
      }
      throw SpherocylError(st, msg);
    }
    m_sys = s;
    check(sc_set_precision(m_sys, m_nPrecision), "set_precision");
    check(sc_set_contact_model(m_sys, m_nContactModel), "set_contact_model");
    m_bGeomStale = false;
    m_bDevStale = true;
  }
  if (m_bDevStale) {
    check(sc_set_positions(m_sys, m_vX.data(), m_vY.data(), m_vPhi.data()), "set_positions");
    check(sc_set_gamma(m_sys, 0, m_dGamma, m_dGammaTotal), "set_gamma");
    m_bDevStale = false;
  }
  check(sc_set_dynamics(m_sys, 0, m_dStrainRate, m_dResizeRate, m_dStep), "set_dynamics");
}

void SpherocylBox::read_scalars() {
  double b[SC_BOX_NFIELDS];
  check(sc_get_box(m_sys, 0, b), "get_box");
  m_dLx = b[SC_BOX_LX];
  m_dLy = b[SC_BOX_LY];
  m_dGamma = b[SC_BOX_GAMMA];
  m_dGammaTotal = b[SC_BOX_GAMMA_TOTAL];
  m_dFrms = b[SC_BOX_FRMS];
  m_nContacts = (int)b[SC_BOX_NCONTACTS];
}

void SpherocylBox::sync_host() {
  if (!m_sys || !m_bHostStale) return;
  check(sc_get_positions(m_sys, m_vX.data(), m_vY.data(), m_vPhi.data()), "get_positions");
  read_scalars();
  m_bHostStale = false;
}

sc_system* SpherocylBox::handle() {
  ensure_device();
  m_bHostStale = true;  // the caller may advance the system directly
  return m_sys;
}

void SpherocylBox::configure_cells() { ensure_device(); }
void SpherocylBox::find_cells() { ensure_device(); }
void SpherocylBox::find_neighbors() { ensure_device(); }

// ------------------------------------------------------------ force passes

void SpherocylBox::calc_forces() {
  ensure_device();
  check(sc_calc_forces(m_sys), "calc_forces");
  read_scalars();
}

void SpherocylBox::calc_se() {
  ensure_device();
  check(sc_calc_se(m_sys), "calc_se");
  double b[SC_BOX_NFIELDS];
  check(sc_get_box(m_sys, 0, b), "get_box");
  m_dEnergy = b[SC_BOX_E];
  m_dPxx = b[SC_BOX_PXX];
  m_dPyy = b[SC_BOX_PYY];
  m_dPxy = b[SC_BOX_PXY];
  for (int i = 0; i < 4; ++i) m_dCorr[i] = b[SC_BOX_CPXX + i];
  read_scalars();
}

void SpherocylBox::stress(double* pdPxx, double* pdPyy, double* pdPxy) const {
  if (pdPxx) *pdPxx = m_dPxx;
  if (pdPyy) *pdPyy = m_dPyy;
  if (pdPxy) *pdPxy = m_dPxy;
}

void SpherocylBox::corrected_stress(double* pdP) const {
  for (int i = 0; i < 4; ++i) pdP[i] = m_dCorr[i];
}

double SpherocylBox::residual_force() {
  if (m_sys) read_scalars();
  return m_dFrms;
}

double SpherocylBox::gamma() {
  sync_host();
  return m_dGamma;
}

double SpherocylBox::total_gamma() {
  sync_host();
  return m_dGammaTotal;
}

void SpherocylBox::positions(std::vector<double>& x, std::vector<double>& y,
                             std::vector<double>& phi) {
  sync_host();
  x = m_vX;
  y = m_vY;
  phi = m_vPhi;
}

void SpherocylBox::forces(std::vector<double>& fx, std::vector<double>& fy,
                          std::vector<double>& tau) {
  ensure_device();
  fx.resize(m_nParticles);
  fy.resize(m_nParticles);
  tau.resize(m_nParticles);
  check(sc_get_forces(m_sys, fx.data(), fy.data(), tau.data()), "get_forces");
}

static std::vector<std::pair<int, int>> pairs_of(
    sc_system* s, int (*fn)(sc_system*, int, int*, int*, int, int*)) {
  int n = 0;
  int st = fn(s, 0, nullptr, nullptr, 0, &n);
  if (st != SC_OK) throw SpherocylError(st, sc_last_error(s));
  std::vector<int> p(std::max(n, 1)), q(std::max(n, 1));
  st = fn(s, 0, p.data(), q.data(), n, &n);
  if (st != SC_OK) throw SpherocylError(st, sc_last_error(s));
  std::vector<std::pair<int, int>> out(n);
  for (int i = 0; i < n; ++i) out[i] = {p[i], q[i]};
  return out;
}

std::vector<std::pair<int, int>> SpherocylBox::contacts() {
  ensure_device();
  return pairs_of(m_sys, sc_get_contacts);
}

std::vector<std::pair<int, int>> SpherocylBox::neighbors() {
  ensure_device();
  return pairs_of(m_sys, sc_get_neighbors);
}

// ------------------------------------------------------------ output

void SpherocylBox::save_positions(std::string strFile) {
  sync_host();
  std::string path = m_strOutputDir + "/" + strFile;
  if (sc_write_positions(path.c_str(), m_nParticles, m_dLx, m_dPacking, m_dGamma, m_dGammaTotal,
                         m_dStrainRate, m_dStep, m_vX.data(), m_vY.data(), m_vPhi.data(),
                         m_vR.data(), m_vA.data()) != SC_OK)
    throw SpherocylError(SC_ERR_ARG, "save_positions: cannot write " + path);
}

void SpherocylBox::write_se_line(std::ostream& out, unsigned long t, bool bPacking) {
  double v[5];
  int k = 0;
  if (bPacking) v[k++] = m_dPacking;
  v[k++] = m_dEnergy;
  v[k++] = m_dPxx;
  v[k++] = m_dPyy;
  v[k++] = m_dPxy;
  se_line(out, t, k, v);
}

void SpherocylBox::display(bool bParticles, bool bCells) {
  sync_host();
  std::cout << "Dimensions: " << m_dLx << " x " << m_dLy << std::endl;
  std::cout << "Particles: " << m_nParticles << std::endl;
  std::cout << "Packing fraction: " << m_dPacking << std::endl;
  std::cout << "Gamma: " << m_dGamma << " (" << m_dGammaTotal << ")" << std::endl;
  std::cout << "Strain rate: " << m_dStrainRate << std::endl;
  std::cout << "Integration step size: " << m_dStep << std::endl;
  std::cout << "\nEnergy: " << m_dEnergy << std::endl;
  std::cout << "Pxx: " << m_dPxx << std::endl;
  std::cout << "Pyy: " << m_dPyy << std::endl;
  std::cout << "Pxy: " << m_dPxy << std::endl;
  if (bParticles && m_nParticles > 0) {
    std::vector<double> fx, fy, tau;
    forces(fx, fy, tau);
    std::cout << "\nParticles: (x, y, phi), (A, R) - (Fx, Fy, Tau)" << std::endl;
    for (int p = 0; p < m_nParticles; ++p)
      std::cout << p << ": (" << m_vX[p] << ", " << m_vY[p] << ", " << m_vPhi[p] << "), ("
                << m_vA[p] << ", " << m_vR[p] << ") - (" << fx[p] << ", " << fy[p] << ", "
                << tau[p] << ")\n";
  }
  if (bCells && m_sys) {
    double b[SC_BOX_NFIELDS];
    check(sc_get_box(m_sys, 0, b), "get_box");
    std::cout << "\nCells: " << (int)b[SC_BOX_ROWS] << " rows x " << (int)b[SC_BOX_COLS]
              << " columns, padding " << b[SC_BOX_PAD] << std::endl;
  }
}

// ------------------------------------------------------------ configurations

void SpherocylBox::random_configuration(double dA, double dR, initialConfig config) {
  if (!(dA >= 0) || !(dR > 0)) throw SpherocylError(SC_ERR_ARG, "random_configuration: A < 0 or R <= 0");
  if (m_nParticles <= 0) throw SpherocylError(SC_ERR_STATE, "random_configuration: no particles");
  std::vector<double> A(m_nParticles, dA), R(m_nParticles, dR);
  std::vector<double> x(m_nParticles), y(m_nParticles), phi(m_nParticles);
  int st = sc_random_configuration(m_nParticles, m_dLx, m_dGamma, A.data(), R.data(), (int)config,
                                   m_nSeed, 100000, x.data(), y.data(), phi.data());
  if (st != SC_OK)
    throw SpherocylError(st, std::string("random_configuration: ") + sc_status_string(st));
  sync_host();
  m_vX = x;
  m_vY = y;
  m_vPhi = phi;
  m_vA = A;
  m_vR = R;
  m_dPacking = packing_from_arrays();
  m_bGeomStale = true;
  m_bDevStale = true;
}

// ------------------------------------------------------------ protocols

long unsigned int SpherocylBox::run_strain(double dRunLength, long unsigned int nTime) {
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  std::fstream outfSE;
  if (nTime == 0) {
    outfSE.open(sePath.c_str(), std::ios::out);
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "run_strain: cannot open " + sePath);
    calc_se();
    outfSE << 0 << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " " << m_dPxy
           << std::endl;
    save_positions("sp0000000000.dat");
  } else {
    outfSE.open(sePath.c_str(), std::ios::in | std::ios::out);
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "run_strain: cannot open " + sePath);
    if (!seek_after_time(outfSE, nTime))
      throw SpherocylError(SC_ERR_ARG, "run_strain: start time not found in " + sePath);
    calc_forces();
  }
  if (!(m_dStrainRate > 0) || !(m_dStep > 0))
    throw SpherocylError(SC_ERR_ARG, "run_strain: strain rate and step must be positive");
  const long nTSteps = long(dRunLength / m_dStrainRate + 0.5);
  const long nIntSteps = long(1.0 / m_dStep + 0.5);
  const long nPosSaveT = long(m_dPosSaveStep / m_dStrainRate + 0.5);
  const long nSESaveT = long(m_dSESaveStep / (m_dStrainRate * m_dStep) + 0.5);
  if (nIntSteps < 1 || nPosSaveT < 1 || nSESaveT < 1)
    throw SpherocylError(SC_ERR_ARG,
                         "run_strain: step, position or SE save cadence rounds to zero");
  ensure_device();
  // Outer steps are batched into one GPU call up to the next position save when
  // the SE cadence divides the sub-steps per outer step (then the records fall
  // on the same sub-steps as the reference's per-outer-step counter).
  const bool batch = nIntSteps % nSESaveT == 0;
  const int seEvery = nSESaveT <= nIntSteps ? (int)nSESaveT : 0;
  std::vector<double> rec;
  unsigned long t = nTime + 1;
  const unsigned long tEnd = nTime + nTSteps;
  while (t <= tEnd) {
    unsigned long m = 1;
    if (batch) {
      unsigned long nextSave = ((t + nPosSaveT - 1) / nPosSaveT) * nPosSaveT;
      m = std::min(nextSave, tEnd) - t + 1;
      long perStep = nIntSteps / nSESaveT;
      m = std::min<unsigned long>(m, std::max<long>(1, kMaxRecordsPerCall / std::max(perStep, 1L)));
    }
    const long nsub = (long)m * nIntSteps;
    if (nsub > std::numeric_limits<int>::max())
      throw SpherocylError(SC_ERR_ARG, "run_strain: too many sub-steps per outer step");
    const long nrec = seEvery > 0 ? nsub / seEvery : 0;
    rec.resize(std::max<long>(nrec, 1) * SC_SE_NFIELDS);
    check(sc_strain_steps(m_sys, (int)nsub, seEvery, rec.data()), "strain_steps");
    m_bHostStale = true;
    long r = 0;
    for (unsigned long k = 0; k < m; ++k, ++t) {
      // records of this outer step: global sub-steps (r+1)*seEvery within it
      while (r < nrec && (unsigned long)(((r + 1) * seEvery - 1) / nIntSteps) == k) {
        const double* v = &rec[r * SC_SE_NFIELDS];
        m_dEnergy = v[0];
        m_dPxx = v[1];
        m_dPyy = v[2];
        m_dPxy = v[3];
        for (int i = 0; i < 4; ++i) m_dCorr[i] = v[4 + i];
        write_se_line(outfSE, t, false);
        ++r;
      }
      if ((t * (unsigned long)nIntSteps) % (unsigned long)nSESaveT == 0)
        write_se_line(outfSE, t, false);
      if (t % (unsigned long)nPosSaveT == 0) {
        outfSE.flush();
        save_positions(time_name("sp", t));
      }
    }
  }
  read_scalars();
  outfSE.flush();
  outfSE.close();
  nTime += nTSteps;
  save_positions(time_name("sp", nTime));
  return nTime;
}

long unsigned int SpherocylBox::run_strain(double dRunLength) {
  // Continues from the last time stamp of the SE file (0 if there is none);
  // src/cpp/strain_dynamics.cpp:321-393.  Lines carry Pyx as a sixth value
  // (the corrected virial component: the reference's own never resets).
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  unsigned long nTime = 0;
  {
    std::ifstream inf(sePath.c_str());
    std::string line, last;
    while (std::getline(inf, line))
      if (!line.empty()) last = line;
    if (!last.empty()) nTime = std::strtoul(last.substr(0, last.find(' ')).c_str(), nullptr, 10);
  }
  std::ofstream outfSE;
  outfSE.precision(14);
  if (nTime == 0) {
    outfSE.open(sePath.c_str());
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "run_strain: cannot open " + sePath);
    calc_se();
    outfSE << 0 << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " " << m_dPxy << " "
           << m_dCorr[3] << std::endl;
    save_positions("sp0000000000.dat");
  } else {
    outfSE.open(sePath.c_str(), std::ios::app);
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "run_strain: cannot open " + sePath);
    calc_forces();
  }
  if (!(m_dStrainRate > 0) || !(m_dStep > 0))
    throw SpherocylError(SC_ERR_ARG, "run_strain: strain rate and step must be positive");
  const long nTSteps = long(dRunLength / m_dStrainRate + 0.5);
  const long nIntSteps = long(1.0 / m_dStep + 0.5);
  const long nPosSaveT = long(m_dPosSaveStep / m_dStrainRate + 0.5);
  const long nSESaveT = long(m_dSESaveStep / (m_dStrainRate * m_dStep) + 0.5);
  if (nIntSteps < 1 || nPosSaveT < 1 || nSESaveT < 1)
    throw SpherocylError(SC_ERR_ARG,
                         "run_strain: step, position or SE save cadence rounds to zero");
  ensure_device();
  const int seEvery = nSESaveT <= nIntSteps ? (int)nSESaveT : 0;
  const long nrec = seEvery > 0 ? nIntSteps / seEvery : 0;
  std::vector<double> rec(std::max<long>(nrec, 1) * SC_SE_NFIELDS);
  auto line = [&](unsigned long t) {
    outfSE << t << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " " << m_dPxy << " "
           << m_dCorr[3] << "\n";
  };
  for (unsigned long t = nTime + 1; t <= nTime + nTSteps; ++t) {
    check(sc_strain_steps(m_sys, (int)nIntSteps, seEvery, rec.data()), "strain_steps");
    m_bHostStale = true;
    for (long r = 0; r < nrec; ++r) {
      const double* v = &rec[r * SC_SE_NFIELDS];
      m_dEnergy = v[0];
      m_dPxx = v[1];
      m_dPyy = v[2];
      m_dPxy = v[3];
      for (int i = 0; i < 4; ++i) m_dCorr[i] = v[4 + i];
      line(t);
    }
    if ((t * (unsigned long)nIntSteps) % (unsigned long)nSESaveT == 0) line(t);
    if (t % (unsigned long)nPosSaveT == 0) {
      outfSE.flush();
      save_positions(time_name("sp", t));
    }
  }
  read_scalars();
  outfSE.flush();
  outfSE.close();
  nTime += nTSteps;
  save_positions(time_name("sp", nTime));
  return nTime;
}

long unsigned int SpherocylBox::resize_box(double dFinalPacking, long unsigned int nTime) {
  if (!(m_dResizeRate != 0.0) || !(m_dStep > 0))
    throw SpherocylError(SC_ERR_ARG, "resize_box: resize rate must be non-zero, step positive");
  sync_host();
  m_dPacking = packing_from_arrays();
  const double dPArea = m_dPacking * (m_dLx * m_dLy);
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  std::fstream outfSE;
  if (nTime == 0) {
    outfSE.open(sePath.c_str(), std::ios::out);
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "resize_box: cannot open " + sePath);
    calc_se();
    outfSE << 0 << " " << m_dPacking << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " "
           << m_dPxy << std::endl;
    save_positions("sp0000000000.dat");
  } else {
    outfSE.open(sePath.c_str(), std::ios::in | std::ios::out);
    if (!outfSE) throw SpherocylError(SC_ERR_ARG, "resize_box: cannot open " + sePath);
    if (!seek_after_time(outfSE, nTime))
      throw SpherocylError(SC_ERR_ARG, "resize_box: start time not found in " + sePath);
    calc_forces();
  }
  const long nIntSteps = long(1.0 / m_dStep + 0.5);
  const long nPosSaveT = long(m_dPosSaveStep / m_dResizeRate + 0.5);
  const long nSESaveT = long(m_dSESaveStep / (m_dResizeRate * m_dStep) + 0.5);
  if (nIntSteps < 1 || nPosSaveT < 1 || nSESaveT < 1)
    throw SpherocylError(SC_ERR_ARG,
                         "resize_box: step, position or SE save cadence rounds to zero");
  ensure_device();
  const bool bShrink = dFinalPacking > m_dPacking;
  const int seEvery = nSESaveT <= nIntSteps ? (int)nSESaveT : 0;
  const long nrec = seEvery > 0 ? nIntSteps / seEvery : 0;
  std::vector<double> rec(std::max<long>(nrec, 1) * SC_SE_NFIELDS);
  // the box length after each sub-step, with the device's arithmetic
  const double scale = std::pow(1.0 - m_dResizeRate, m_dStep);
  while (bShrink ? dFinalPacking > m_dPacking : dFinalPacking < m_dPacking) {
    nTime += 1;
    check(sc_resize_steps(m_sys, (int)nIntSteps, bShrink ? 1 : 0, m_bResizeCorrected ? 1 : 0,
                          seEvery, rec.data()),
          "resize_steps");
    m_bHostStale = true;
    double L = m_dLx;
    long r = 0;
    for (long s = 1; s <= nIntSteps; ++s) {
      L = bShrink ? L * scale : L / scale;
      if (seEvery > 0 && s % seEvery == 0 && r < nrec) {
        const double* v = &rec[r * SC_SE_NFIELDS];
        m_dEnergy = v[0];
        m_dPxx = v[1];
        m_dPyy = v[2];
        m_dPxy = v[3];
        for (int i = 0; i < 4; ++i) m_dCorr[i] = v[4 + i];
        m_dPacking = dPArea / (L * L);
        write_se_line(outfSE, nTime, true);
        ++r;
      }
    }
    read_scalars();
    m_dPacking = dPArea / (m_dLx * m_dLy);
    if ((nTime * (unsigned long)nIntSteps) % (unsigned long)nSESaveT == 0)
      write_se_line(outfSE, nTime, true);
    if (nTime % (unsigned long)nPosSaveT == 0) {
      outfSE.flush();
      save_positions(time_name("sp", nTime));
    }
  }
  outfSE.flush();
  outfSE.close();
  save_positions(time_name("sp", nTime));
  return nTime;
}

long unsigned int SpherocylBox::resize_box(double dFinalPacking) {
  // Continues from the last time stamp of the SE file, appending with 14
  // digits (src/cpp/resize_dynamics.cpp:331-416).
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  unsigned long nTime = 0;
  {
    std::ifstream inf(sePath.c_str());
    std::string line, last;
    while (std::getline(inf, line))
      if (!line.empty()) last = line;
    if (!last.empty()) nTime = std::strtoul(last.substr(0, last.find(' ')).c_str(), nullptr, 10);
  }
  if (nTime == 0) return resize_box(dFinalPacking, 0UL);
  if (!(m_dResizeRate != 0.0) || !(m_dStep > 0))
    throw SpherocylError(SC_ERR_ARG, "resize_box: resize rate must be non-zero, step positive");
  sync_host();
  m_dPacking = packing_from_arrays();
  const double dPArea = m_dPacking * (m_dLx * m_dLy);
  std::ofstream outfSE(sePath.c_str(), std::ios::app);
  if (!outfSE) throw SpherocylError(SC_ERR_ARG, "resize_box: cannot open " + sePath);
  outfSE.precision(14);
  calc_forces();
  const long nIntSteps = long(1.0 / m_dStep + 0.5);
  const long nPosSaveT = long(m_dPosSaveStep / m_dResizeRate + 0.5);
  const long nSESaveT = long(m_dSESaveStep / (m_dResizeRate * m_dStep) + 0.5);
  if (nIntSteps < 1 || nPosSaveT < 1 || nSESaveT < 1)
    throw SpherocylError(SC_ERR_ARG,
                         "resize_box: step, position or SE save cadence rounds to zero");
  ensure_device();
  const bool bShrink = dFinalPacking > m_dPacking;
  const int seEvery = nSESaveT <= nIntSteps ? (int)nSESaveT : 0;
  const long nrec = seEvery > 0 ? nIntSteps / seEvery : 0;
  std::vector<double> rec(std::max<long>(nrec, 1) * SC_SE_NFIELDS);
  const double scale = std::pow(1.0 - m_dResizeRate, m_dStep);
  while (bShrink ? dFinalPacking > m_dPacking : dFinalPacking < m_dPacking) {
    nTime += 1;
    check(sc_resize_steps(m_sys, (int)nIntSteps, bShrink ? 1 : 0, m_bResizeCorrected ? 1 : 0,
                          seEvery, rec.data()),
          "resize_steps");
    m_bHostStale = true;
    double L = m_dLx;
    long r = 0;
    for (long s = 1; s <= nIntSteps; ++s) {
      L = bShrink ? L * scale : L / scale;
      if (seEvery > 0 && s % seEvery == 0 && r < nrec) {
        const double* v = &rec[r * SC_SE_NFIELDS];
        m_dEnergy = v[0];
        m_dPxx = v[1];
        m_dPyy = v[2];
        m_dPxy = v[3];
        for (int i = 0; i < 4; ++i) m_dCorr[i] = v[4 + i];
        m_dPacking = dPArea / (L * L);
        write_se_line(outfSE, nTime, true);
        ++r;
      }
    }
    read_scalars();
    m_dPacking = dPArea / (m_dLx * m_dLy);
    if ((nTime * (unsigned long)nIntSteps) % (unsigned long)nSESaveT == 0)
      write_se_line(outfSE, nTime, true);
    if (nTime % (unsigned long)nPosSaveT == 0) {
      outfSE.flush();
      save_positions(time_name("sp", nTime));
    }
  }
  outfSE.flush();
  outfSE.close();
  save_positions(time_name("sp", nTime));
  return nTime;
}

void SpherocylBox::gradient_descent_minimize(int nMaxSteps, double dMinE) {
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  std::fstream outfSE(sePath.c_str(), std::ios::out);
  if (!outfSE) throw SpherocylError(SC_ERR_ARG, "gradient_descent_minimize: cannot open " + sePath);
  const long nPosSaveT = long(m_dPosSaveStep + 0.5);
  if (nPosSaveT < 1)
    throw SpherocylError(SC_ERR_ARG, "gradient_descent_minimize: position save cadence < 1");
  ensure_device();
  calc_se();
  bool hitMin = false;
  for (int s = 0; s < nMaxSteps; ++s) {
    outfSE << s << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " " << m_dPxy
           << std::endl;
    if (m_dEnergy <= dMinE) {
      outfSE.flush();
      save_positions(time_name("sr", (unsigned long)s));
      hitMin = true;
      break;
    } else if (s % nPosSaveT == 0) {
      outfSE.flush();
      save_positions(time_name("sr", (unsigned long)s));
    }
    // one step x += dt F, phi += dt tau / I, then forces, energy and stress of
    // the new state (the next iteration's calc_se)
    int done = 0;
    check(sc_gd_steps(m_sys, 1, -std::numeric_limits<double>::infinity(), &done), "gd_steps");
    m_bHostStale = true;
    double b[SC_BOX_NFIELDS];
    check(sc_get_box(m_sys, 0, b), "get_box");
    m_dEnergy = b[SC_BOX_E];
    m_dPxx = b[SC_BOX_PXX];
    m_dPyy = b[SC_BOX_PYY];
    m_dPxy = b[SC_BOX_PXY];
    for (int i = 0; i < 4; ++i) m_dCorr[i] = b[SC_BOX_CPXX + i];
  }
  if (!hitMin) {
    std::cout << "\nMaximum relaxation steps reached" << std::endl;
    outfSE << nMaxSteps << " " << m_dEnergy << " " << m_dPxx << " " << m_dPyy << " " << m_dPxy
           << std::endl;
    save_positions(time_name("sr", (unsigned long)nMaxSteps));
  } else {
    std::cout << "\nEncountered zero energy" << std::endl;
  }
  read_scalars();
  outfSE.close();
}

int SpherocylBox::fire_minimize(const double* pdParams, bool bThrowIfNotConverged) {
  static const double kDefaults[10] = {0.05, 0.5, 5, 1.1, 0.5, 0.1, 0.99, 0.02, 1e-10, 200000};
  ensure_device();
  const double* p = pdParams ? pdParams : kDefaults;
  int iters = 0;
  double frms = 0;
  // mixed precision: FP32 pair math down to a residual of 1e-6, FP64 after
  const double sw = m_nPrecision == SC_PREC_MIXED ? 1e-6 : 0.0;
  int st = sc_fire_minimize(m_sys, p, sw, &iters, &frms);
  m_bHostStale = true;
  if (st == SC_ERR_NOT_CONVERGED && !bThrowIfNotConverged) st = SC_OK;
  check(st, "fire_minimize");
  double b[SC_BOX_NFIELDS];
  check(sc_get_box(m_sys, 0, b), "get_box");
  m_dEnergy = b[SC_BOX_E];
  m_dPxx = b[SC_BOX_PXX];
  m_dPyy = b[SC_BOX_PYY];
  m_dPxy = b[SC_BOX_PXY];
  for (int i = 0; i < 4; ++i) m_dCorr[i] = b[SC_BOX_CPXX + i];
  read_scalars();
  return iters;
}

// ------------------------------------------------------------ quasi-static protocols

void SpherocylBox::save_positions_sd(unsigned long nTime) {
  sync_host();
  std::string path = m_strOutputDir + "/" + time_name("sd", nTime);
  if (sc_write_positions_sd(path.c_str(), m_nParticles, m_dLx, m_dLy, m_dPacking, m_dGamma,
                            m_dGammaTotal, m_dStrainRate, m_dStep, m_vX.data(), m_vY.data(),
                            m_vPhi.data(), m_vR.data(), m_vA.data()) != SC_OK)
    throw SpherocylError(SC_ERR_ARG, "save_positions_sd: cannot write " + path);
}

void SpherocylBox::save_positions_bin(unsigned long nTime) {
  sync_host();
  char name[32];
  std::snprintf(name, sizeof(name), "sd%010lu.bin", nTime);
  std::string path = m_strOutputDir + "/" + name;
  if (sc_write_positions_bin(path.c_str(), m_nParticles, m_dLx, m_dLy, m_dPacking, m_dGamma,
                             m_dGammaTotal, m_dStrainRate, m_dStep, m_vX.data(), m_vY.data(),
                             m_vPhi.data(), m_vR.data(), m_vA.data()) != SC_OK)
    throw SpherocylError(SC_ERR_ARG, "save_positions_bin: cannot write " + path);
}

// FIRE to the box's tolerance; energy, stress, box length and packing
// fraction of the minimised state.  A minimisation that stops at max_iter
// is kept (the reference's minimiser also stops at its step limit) and shows
// in qs_iterations.
int SpherocylBox::qs_minimize() {
  int it = fire_minimize(m_dFire, false);
  m_dPacking = packing_from_arrays();
  m_vQsIters.push_back(it);
  return it;
}

// One quasi-static step as the protocols take it: m_dStep = 0.01, then
// resize_step (src/cuda/strain_dynamics.cu:1124-1196), resize_step_2p
// (src/cuda/2point_dynamics.cu:1187-1285) or its y-only form y_resize_step_2p
// (:1287-1385) -- x, y, L *= 1 + eps (y and Ly only for the last), where the
// two-point form expands by 1/(1 - eps) so that an expansion undoes a
// compression by the same step, followed by one unsheared Heun step of
// dt 0.01 -- then the minimiser.  The box keeps the step 0.01 afterwards, as
// the reference's does.
void SpherocylBox::qs_step(double dEpsilon, bool bTwoPoint, bool bYOnly) {
  ensure_device();
  if (bTwoPoint && dEpsilon > 0) dEpsilon = 1.0 / (1.0 - dEpsilon) - 1.0;
  const double scale = 1.0 + dEpsilon;
  if (bYOnly)
    check(sc_affine_resize_y(m_sys, &scale), "affine_resize_y");
  else
    check(sc_affine_resize(m_sys, &scale), "affine_resize");
  m_dStep = 0.01;
  check(sc_set_dynamics(m_sys, 0, 0.0, 0.0, m_dStep), "set_dynamics");
  check(sc_strain_steps(m_sys, 1, 0, nullptr), "strain_steps");
  check(sc_set_dynamics(m_sys, 0, m_dStrainRate, m_dResizeRate, m_dStep), "set_dynamics");
  m_bHostStale = true;
  read_scalars();
  qs_minimize();
}

void SpherocylBox::qs_line(FILE* f, const char* fmt, unsigned long t) {
  std::fprintf(f, fmt, (int)t, m_dPacking, m_dEnergy, m_dCorr[0], m_dCorr[1], m_dCorr[2],
               m_dCorr[3]);
  std::fflush(f);
}

// ------------------------------------------------------------ pure shear

void SpherocylBox::flip_shear_direction() {
  ensure_device();
  check(sc_flip_shear(m_sys, 0), "flip_shear");
  m_bHostStale = true;
  read_scalars();
}

void SpherocylBox::rotate_by_gamma() {
  ensure_device();
  check(sc_rotate_by_gamma(m_sys, 0), "rotate_by_gamma");
  m_bHostStale = true;
  read_scalars();
}

long unsigned int SpherocylBox::run_pure_strain(long unsigned int nStart, double dStopAspect,
                                                double dSvStressGamma, double dSvPosGamma) {
  if (m_dStrainRate == 0.0) throw SpherocylError(SC_ERR_ARG, "run_pure_strain: zero strain rate");
  m_dStrainRate = std::fabs(m_dStrainRate);  // as the reference does (2point_dynamics.cu:2292-2295)
  ensure_device();
  read_scalars();
  if (m_dLx / m_dLy > dStopAspect)
    throw SpherocylError(SC_ERR_ARG, "run_pure_strain: box aspect already beyond the stop");
  return run_sd(nStart, true, 0, dStopAspect, dSvStressGamma, dSvPosGamma, "run_pure_strain");
}

// run_strain of the GPU class (src/cuda/strain_dynamics.cu:984-1083)
long unsigned int SpherocylBox::run_strain_sd(long unsigned int nStart, double dRunGamma,
                                              double dSvStressGamma, double dSvPosGamma) {
  if (!(m_dStrainRate > 0)) throw SpherocylError(SC_ERR_ARG, "run_strain_sd: strain rate must be positive");
  if (!(dRunGamma >= 0) || !std::isfinite(dRunGamma))
    throw SpherocylError(SC_ERR_ARG, "run_strain_sd: run length must be >= 0");
  ensure_device();
  read_scalars();
  const unsigned long nStop = nStart + (unsigned long)(dRunGamma / m_dStrainRate + 0.5);
  return run_sd(nStart, false, nStop, 0.0, dSvStressGamma, dSvPosGamma, "run_strain_sd");
}

// The time loop the GPU class's run_strain and run_pure_strain_2p share: per
// time unit 1/step sub-steps; every nSvStress sub-steps a line "t E Pxx Pyy P
// Pxy Pyx" (printf %.7g; the corrected stress with the CPU code's sign) of the
// state the sub-step starts from, every nSvPos time units also a position file
// sd%010lu.dat; a restart seeks the SE file to nStart.  Runs until nStop
// (sheared) or until Lx / Ly reaches dStopAspect (pure shear).
long unsigned int SpherocylBox::run_sd(long unsigned int nStart, bool bPure, unsigned long nStop,
                                       double dStopAspect, double dSvStressGamma,
                                       double dSvPosGamma, const char* name) {
  const double rate = m_dStrainRate;
  if (dSvStressGamma < rate * m_dStep) dSvStressGamma = rate * m_dStep;
  if (dSvPosGamma < rate) dSvPosGamma = rate;
  const unsigned long nIntStep = (unsigned long)(1.0 / m_dStep + 0.5);
  const unsigned long nSvStress = (unsigned long)(dSvStressGamma / (rate * m_dStep) + 0.5);
  const unsigned long nSvPos = (unsigned long)(dSvPosGamma / rate + 0.5);
  if (nIntStep < 1 || nSvStress < 1 || nSvPos < 1)
    throw SpherocylError(SC_ERR_ARG, std::string(name) + ": step or save cadence rounds to zero");
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  std::fstream outfSE;
  if (nStart == 0) {
    outfSE.open(sePath.c_str(), std::ios::out);
  } else {
    outfSE.open(sePath.c_str(), std::ios::in | std::ios::out);
    if (outfSE && !seek_after_time(outfSE, nStart))
      throw SpherocylError(SC_ERR_ARG, std::string(name) + ": start time not found in " + sePath);
  }
  if (!outfSE) throw SpherocylError(SC_ERR_ARG, std::string(name) + ": cannot open " + sePath);
  if (bPure) check(sc_set_path(m_sys, SC_PATH_LIST), "set_path");
  check(sc_set_dynamics(m_sys, 0, rate, 0.0, m_dStep), "set_dynamics");
  auto line = [&](unsigned long t) {
    calc_se();
    char buf[256];
    std::snprintf(buf, sizeof buf, "%lu %.7g %.7g %.7g %.7g %.7g %.7g\n", t, m_dEnergy,
                  m_dCorr[0], m_dCorr[1], 0.5 * (m_dCorr[0] + m_dCorr[1]), m_dCorr[2], m_dCorr[3]);
    outfSE << buf;
  };
  auto more = [&](unsigned long t) { return bPure ? m_dLx / m_dLy < dStopAspect : t < nStop; };
  unsigned long nTime = nStart, nTotal = nTime * nIntStep;
  while (more(nTime)) {
    unsigned long i = 0;
    while (i < nIntStep) {
      const bool svPos = i == 0 && nTime % nSvPos == 0;
      if (svPos || nTotal % nSvStress == 0) {
        line(nTime);
        if (svPos) {
          outfSE.flush();
          save_positions_sd(nTime);
        }
      }
      // steps up to the next stress record or the end of this time unit
      unsigned long k = std::min(nIntStep - i, nSvStress - nTotal % nSvStress);
      if (bPure)
        check(sc_pure_strain_steps(m_sys, (int)k, 0, nullptr), "pure_strain_steps");
      else
        check(sc_strain_steps(m_sys, (int)k, 0, nullptr), "strain_steps");
      m_bHostStale = true;
      i += k;
      nTotal += k;
    }
    ++nTime;
    read_scalars();
  }
  line(nTime);
  outfSE.flush();
  save_positions_sd(nTime);
  outfSE.close();
  return nTime;
}

// compress_qs2p / expand_qs2p (src/cuda/2point_dynamics.cu:4300-4390)
long unsigned int SpherocylBox::qs_run(double dLimit, double dResizeStep, bool bCompress) {
  if (dResizeStep == 0 || !std::isfinite(dResizeStep))
    throw SpherocylError(SC_ERR_ARG, "quasi-static resize step must be nonzero");
  if (m_nQsPosEvery < 1) throw SpherocylError(SC_ERR_ARG, "position interval < 1");
  dResizeStep = std::fabs(dResizeStep);
  ensure_device();
  m_dPacking = packing_from_arrays();
  if (bCompress ? !(dLimit >= m_dPacking) : !(dLimit <= m_dPacking))
    throw SpherocylError(SC_ERR_ARG, "quasi-static target packing is on the wrong side");
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  FILE* f = std::fopen(sePath.c_str(), "w");
  if (!f) throw SpherocylError(SC_ERR_ARG, "quasi-static protocol: cannot open " + sePath);
  m_vQsIters.clear();
  unsigned long t = 0;
  try {
    qs_minimize();
    save_positions_sd(t);
    qs_line(f, "%d %f %g %g %g %g %g\n", t);
    while (bCompress ? m_dPacking < dLimit : m_dPacking > dLimit) {
      ++t;
      qs_step(bCompress ? -dResizeStep : dResizeStep, true);
      if (t % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(t);
      qs_line(f, "%d %.9f %g %g %g %g %g\n", t);
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return t;
}

long unsigned int SpherocylBox::compress_qs2p(double dMaxPacking, double dResizeStep) {
  return qs_run(dMaxPacking, dResizeStep, true);
}

long unsigned int SpherocylBox::expand_qs2p(double dMinPacking, double dResizeStep) {
  return qs_run(dMinPacking, dResizeStep, false);
}

// resize_to_energy_2p (src/cuda/2point_dynamics.cu:4393-4450): the energy
// is that of the current state (the reference's calc_fe pass).
// y_resize_to_energy_2p (:4458-4520) is the same search with y-only steps.
long unsigned int SpherocylBox::resize_to_energy_2p(FILE* f, long unsigned int nTime,
                                                     double dEnergy, double dStep, bool bYOnly) {
  dStep = std::fabs(dStep);
  calc_se();
  const bool compress = m_dEnergy < dEnergy;
  while (compress ? m_dEnergy < dEnergy : m_dEnergy > dEnergy) {
    ++nTime;
    qs_step(compress ? -dStep : dStep, true, bYOnly);
    if (nTime % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(nTime);
    qs_line(f, "%d %.10f %.9g %.9g %.9g %.9g %.9g\n", nTime);
    dStep *= 1.1;
  }
  return nTime;
}

// find_jam_2p (src/cuda/2point_dynamics.cu:4582-4606)
long unsigned int SpherocylBox::find_jam_2p(double dJamE, double dResizeStep) {
  return jam_search_2p(dJamE, dResizeStep, 1e-15, false, "find_jam_2p");
}

// y_compress_energy_2p (src/cuda/2point_dynamics.cu:4608-4632): the search of
// find_jam_2p with y-only steps and a minimum step of 1e-14
long unsigned int SpherocylBox::y_compress_energy_2p(double dTargetE, double dResizeStep) {
  return jam_search_2p(dTargetE, dResizeStep, 1e-14, true, "y_compress_energy_2p");
}

long unsigned int SpherocylBox::jam_search_2p(double dJamE, double dResizeStep,
                                              double dMinResizeStep, bool bYOnly,
                                              const char* name) {
  if (!(dResizeStep > 0) || !std::isfinite(dResizeStep) || !(dJamE >= 0))
    throw SpherocylError(SC_ERR_ARG, std::string(name) + ": need a target energy >= 0 and a positive step");
  if (m_nQsPosEvery < 1) throw SpherocylError(SC_ERR_ARG, "position interval < 1");
  ensure_device();
  m_dPacking = packing_from_arrays();
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  FILE* f = std::fopen(sePath.c_str(), "w");
  if (!f) throw SpherocylError(SC_ERR_ARG, std::string(name) + ": cannot open " + sePath);
  m_vQsIters.clear();
  unsigned long t = 0, told = 0;
  try {
    save_positions_sd(0);
    while (dResizeStep > dMinResizeStep) {
      t = resize_to_energy_2p(f, t, dJamE, dResizeStep, bYOnly);
      dResizeStep = dResizeStep * std::pow(1.1, (double)((long)t - (long)told - 1)) / 2;
      told = t;
    }
    save_positions_bin(t);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return t;
}

// ------------------------------------------------ one-point quasi-static protocols

FILE* SpherocylBox::qs_open_se() {
  if (m_nQsPosEvery < 1) throw SpherocylError(SC_ERR_ARG, "position interval < 1");
  std::string sePath = m_strOutputDir + "/" + m_strSEOutput;
  FILE* f = std::fopen(sePath.c_str(), "w");
  if (!f) throw SpherocylError(SC_ERR_ARG, "quasi-static protocol: cannot open " + sePath);
  return f;
}

// quasistatic_compress (src/cuda/energy_minimize.cu:1731-1783)
void SpherocylBox::quasistatic_compress(double dMaxPack, double dResizeStep) {
  if (dResizeStep == 0 || !std::isfinite(dResizeStep))
    throw SpherocylError(SC_ERR_ARG, "quasistatic_compress: resize step must be nonzero");
  dResizeStep = std::fabs(dResizeStep);
  ensure_device();
  m_dPacking = packing_from_arrays();
  FILE* f = qs_open_se();
  m_vQsIters.clear();
  unsigned long t = 0;
  try {
    qs_minimize();
    save_positions_sd(t);
    qs_line(f, "%d %f %g %g %g %g %g\n", t);
    while (m_dEnergy > 0) {
      ++t;
      qs_step(dResizeStep, false);
      if (t % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(t);
      qs_line(f, "%d %f %g %g %g %g %g\n", t);
    }
    while (m_dPacking < dMaxPack) {
      ++t;
      qs_step(-dResizeStep, false);
      if (t % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(t);
      qs_line(f, "%d %f %g %g %g %g %g\n", t);
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
}

// qs_one_cycle (src/cuda/energy_minimize.cu:1785-1836): the reference's six
// damped conjugate-gradient passes per step are one FIRE minimisation here
long unsigned int SpherocylBox::qs_one_cycle(FILE* f, long unsigned int nTime, double dMaxE,
                                             double dResizeStep) {
  if (dResizeStep == 0 || !std::isfinite(dResizeStep))
    throw SpherocylError(SC_ERR_ARG, "qs_one_cycle: resize step must be nonzero");
  dResizeStep = std::fabs(dResizeStep);
  ensure_device();
  qs_minimize();
  save_positions_sd(nTime);
  qs_line(f, "%d %f %g %g %g %g %g\n", nTime);
  while (m_dEnergy > 0) {
    ++nTime;
    qs_step(dResizeStep, false);
    if (nTime % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(nTime);
    qs_line(f, "%d %f %g %g %g %g %g\n", nTime);
  }
  while (m_dEnergy < dMaxE) {
    ++nTime;
    qs_step(-dResizeStep, false);
    if (nTime % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(nTime);
    qs_line(f, "%d %f %g %g %g %g %g\n", nTime);
  }
  return nTime;
}

// quasistatic_cycle (src/cuda/energy_minimize.cu:1838-1857)
long unsigned int SpherocylBox::quasistatic_cycle(int nCycles, double dMaxE, double dResizeStep) {
  FILE* f = qs_open_se();
  m_vQsIters.clear();
  unsigned long t = 0;
  try {
    for (int c = 0; c < nCycles; ++c) {
      t = qs_one_cycle(f, t, dMaxE, dResizeStep);
      t += 1;
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return t;
}

// quasistatic_find_jam (src/cuda/energy_minimize.cu:1859-1880)
long unsigned int SpherocylBox::quasistatic_find_jam(double dMaxE, double dMinE,
                                                     double dResizeStep) {
  if (!(dResizeStep != 0) || !std::isfinite(dResizeStep) || !(dMaxE > 0))
    throw SpherocylError(SC_ERR_ARG, "quasistatic_find_jam: need dMaxE > 0 and a nonzero step");
  ensure_device();
  read_scalars();
  FILE* f = qs_open_se();
  m_vQsIters.clear();
  unsigned long t = 0;
  try {
    const double dMinResizeStep = m_dLx * 1e-15;
    dResizeStep = std::fabs(dResizeStep);
    while (dMaxE > dMinE && dResizeStep > dMinResizeStep) {
      t = qs_one_cycle(f, t, dMaxE, dResizeStep);
      dMaxE *= 0.5;
      dResizeStep *= 0.5;
      t += 1;
    }
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return t;
}

// resize_to_energy_1p (src/cuda/energy_minimize.cu:1882-1935): the energy is
// that of the current state (the reference's calc_fe pass)
long unsigned int SpherocylBox::resize_to_energy_1p(FILE* f, long unsigned int nTime,
                                                     double dEnergy, double dStep) {
  dStep = std::fabs(dStep);
  calc_se();
  const bool compress = m_dEnergy < dEnergy;
  while (compress ? m_dEnergy < dEnergy : m_dEnergy > dEnergy) {
    ++nTime;
    qs_step(compress ? -dStep : dStep, false);
    if (nTime % (unsigned long)m_nQsPosEvery == 0) save_positions_sd(nTime);
    qs_line(f, "%d %.9f %g %g %g %g %g\n", nTime);
    dStep *= 1.1;
  }
  return nTime;
}

// find_jam_1p (src/cuda/energy_minimize.cu:1937-1961)
long unsigned int SpherocylBox::find_jam_1p(double dJamE, double dResizeStep) {
  if (!(dResizeStep > 0) || !std::isfinite(dResizeStep) || !(dJamE >= 0))
    throw SpherocylError(SC_ERR_ARG, "find_jam_1p: need dJamE >= 0 and a positive step");
  ensure_device();
  read_scalars();
  m_dPacking = packing_from_arrays();
  FILE* f = qs_open_se();
  m_vQsIters.clear();
  unsigned long t = 0, told = 0;
  try {
    save_positions_sd(0);
    const double dMinResizeStep = m_dLx * 1e-13;
    while (dResizeStep > dMinResizeStep) {
      t = resize_to_energy_1p(f, t, dJamE, dResizeStep);
      dResizeStep = dResizeStep * std::pow(1.1, (double)((long)t - (long)told)) / 2;
      told = t;
    }
    save_positions_bin(t);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  return t;
}

// ------------------------------------------------------------ position files

PositionFile read_position_file(const std::string& path) {
  PositionFile f;
  int n = 0;
  double h[6];
  if (sc_read_positions(path.c_str(), 0, &n, h, nullptr, nullptr, nullptr, nullptr, nullptr) !=
      SC_OK)
    throw SpherocylError(SC_ERR_ARG, "cannot read positions from " + path);
  f.x.resize(n);
  f.y.resize(n);
  f.phi.resize(n);
  f.R.resize(n);
  f.A.resize(n);
  if (sc_read_positions(path.c_str(), n, &f.n, h, f.x.data(), f.y.data(), f.phi.data(),
                        f.R.data(), f.A.data()) != SC_OK)
    throw SpherocylError(SC_ERR_ARG, "cannot read positions from " + path);
  f.L = h[0];
  f.packing = h[1];
  f.gamma = h[2];
  f.gamma_total = h[3];
  f.strain_rate = h[4];
  f.step = h[5];
  return f;
}

unsigned long position_file_time(const std::string& path) {
  if (path.size() < 14) throw SpherocylError(SC_ERR_ARG, "file name too short for a time stamp");
  return std::strtoul(path.substr(path.size() - 14, 10).c_str(), nullptr, 10);
}

// ------------------------------------------------------------ C entry points

static void se_line(std::ostream& out, unsigned long t, int nvals, const double* vals) {
  out << t;
  for (int i = 0; i < nvals; ++i) out << " " << vals[i];
  out << "\n";
}

extern "C" int sc_write_positions(const char* path, int n, double L, double packing, double gamma,
                                  double gamma_total, double strain_rate, double step,
                                  const double* x, const double* y, const double* phi,
                                  const double* R, const double* A) {
  if (!path || n < 0 || (n > 0 && (!x || !y || !phi || !R || !A))) return SC_ERR_ARG;
  std::ofstream outf(path);
  if (!outf) return SC_ERR_ARG;
  outf.precision(14);
  outf << n << " " << L << " " << packing << " " << gamma << " " << gamma_total << " "
       << strain_rate << " " << step << std::endl;
  for (int p = 0; p < n; ++p)
    outf << x[p] << " " << y[p] << " " << phi[p] << " " << R[p] << " " << A[p] << std::endl;
  outf.close();
  return outf ? SC_OK : SC_ERR_ARG;
}

extern "C" int sc_read_positions(const char* path, int cap, int* n, double* header, double* x,
                                 double* y, double* phi, double* R, double* A) {
  if (!path || !n || !header || cap < 0 || (cap > 0 && (!x || !y || !phi || !R || !A)))
    return SC_ERR_ARG;
  std::ifstream in(path);
  if (!in) return SC_ERR_ARG;
  std::string head;
  if (!std::getline(in, head)) return SC_ERR_ARG;
  std::istringstream hs(head);
  int nh = 0;
  if (!(hs >> nh)) return SC_ERR_ARG;
  std::vector<double> hv;
  for (double v; hs >> v;) hv.push_back(v);
  if (hv.size() == 7) {  // the GPU code's "N Lx Ly packing ..." (square boxes only)
    if (hv[0] != hv[1]) return SC_ERR_ARG;
    hv.erase(hv.begin() + 1);
  }
  if (hv.size() != 6) return SC_ERR_ARG;
  for (int i = 0; i < 6; ++i) header[i] = hv[i];
  int rows = 0;
  std::string line;
  while (std::getline(in, line)) {
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream ls(line);
    double v[5];
    for (double& d : v)
      if (!(ls >> d)) return SC_ERR_ARG;
    if (rows < cap) {
      x[rows] = v[0];
      y[rows] = v[1];
      phi[rows] = v[2];
      R[rows] = v[3];
      A[rows] = v[4];
    }
    ++rows;
  }
  *n = rows;
  return SC_OK;
}

extern "C" int sc_write_positions_sd(const char* path, int n, double Lx, double Ly,
                                     double packing, double gamma, double gamma_total,
                                     double strain_rate, double step, const double* x,
                                     const double* y, const double* phi, const double* R,
                                     const double* A) {
  if (!path || n < 0 || (n > 0 && (!x || !y || !phi || !R || !A))) return SC_ERR_ARG;
  FILE* f = std::fopen(path, "w");
  if (!f) return SC_ERR_ARG;
  bool ok = std::fprintf(f, "%d %.17g %.17g %f %.17g %.14g %g %g\n", n, Lx, Ly, packing, gamma,
                         gamma_total, strain_rate, step) > 0;
  for (int p = 0; ok && p < n; ++p)
    ok = std::fprintf(f, "%.17g %.17g %.17g %.2g %.17g\n", x[p], y[p], phi[p], R[p], A[p]) > 0;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SC_OK : SC_ERR_ARG;
}

extern "C" int sc_write_positions_bin(const char* path, int n, double Lx, double Ly,
                                      double packing, double gamma, double gamma_total,
                                      double strain_rate, double step, const double* x,
                                      const double* y, const double* phi, const double* R,
                                      const double* A) {
  if (!path || n < 0 || (n > 0 && (!x || !y || !phi || !R || !A))) return SC_ERR_ARG;
  FILE* f = std::fopen(path, "wb");
  if (!f) return SC_ERR_ARG;
  const double head[7] = {Lx, Ly, packing, gamma, gamma_total, strain_rate, step};
  bool ok = std::fwrite(&n, sizeof(int), 1, f) == 1 && std::fwrite(head, sizeof(double), 7, f) == 7;
  for (int p = 0; ok && p < n; ++p) {
    const double rec[5] = {x[p], y[p], phi[p], R[p], A[p]};
    ok = std::fwrite(rec, sizeof(double), 5, f) == 5;
  }
  ok = (std::fclose(f) == 0) && ok;
  return ok ? SC_OK : SC_ERR_ARG;
}

extern "C" int sc_read_positions_bin(const char* path, int cap, int* n, double* header, double* x,
                                     double* y, double* phi, double* R, double* A) {
  if (!path || !n || !header || cap < 0 || (cap > 0 && (!x || !y || !phi || !R || !A)))
    return SC_ERR_ARG;
  FILE* f = std::fopen(path, "rb");
  if (!f) return SC_ERR_ARG;
  int count = 0;
  bool ok = std::fread(&count, sizeof(int), 1, f) == 1 && count >= 0 &&
            std::fread(header, sizeof(double), 7, f) == 7;
  if (ok) {  // the file must hold exactly count records
    const long body = std::ftell(f);
    ok = std::fseek(f, 0, SEEK_END) == 0 &&
         std::ftell(f) == body + (long)sizeof(double) * 5 * (long)count &&
         std::fseek(f, body, SEEK_SET) == 0;
  }
  for (int p = 0; ok && p < count && p < cap; ++p) {
    double rec[5];
    ok = std::fread(rec, sizeof(double), 5, f) == 5;
    if (ok) {
      x[p] = rec[0];
      y[p] = rec[1];
      phi[p] = rec[2];
      R[p] = rec[3];
      A[p] = rec[4];
    }
  }
  std::fclose(f);
  if (!ok) return SC_ERR_ARG;
  *n = count;
  return SC_OK;
}

extern "C" int sc_format_se_line(char* buf, int cap, unsigned long t, int nvals,
                                 const double* vals, int precision) {
  if (!buf || cap < 1 || nvals < 0 || (nvals > 0 && !vals) || precision < 1) return -1;
  std::ostringstream os;
  os.precision(precision);
  se_line(os, t, nvals, vals);
  const std::string s = os.str();
  if ((int)s.size() + 1 > cap) return -1;
  std::memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}
