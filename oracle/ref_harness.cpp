// Test infrastructure only (never shipped, never on the product path).
//
// C-ABI harness around the reference CPU implementation (src/cpp of
// tmarschall/spherocyls), compiled from the reference's own sources by
// oracle/Makefile into oracle/_ref/libspherocyl_ref.so.  It lets the tests and
// the golden-vector script drive the reference's SpherocylBox directly and read
// the state that the reference only exposes through display():
//   particle records   src/cpp/spherocyl_box.h:10-24
//   box members        src/cpp/spherocyl_box.h:28-64
// The class keeps all data members in one private section, so opening access
// for this translation unit does not change the object layout.
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#define private public
#include "spherocyl_box.h"
#undef private

namespace {
struct Quiet {
  // The reference prints cell geometry to stdout on every configure_cells
  // (src/cpp/spherocyl_box.cpp:234,238); keep the harness silent by default.
  Quiet() { std::cout.setstate(std::ios_base::failbit); }
} g_quiet;
}  // namespace

extern "C" {

void ref_verbose(int on) {
  if (on) std::cout.clear();
  else std::cout.setstate(std::ios_base::failbit);
}

// SpherocylBox(dL, N, pdX, pdY, pdPhi, pdA, pdR, dPadding)
// src/cpp/spherocyl_box.cpp:152-173
void* ref_box_new(double L, int n, const double* x, const double* y, const double* phi,
                  const double* A, const double* R, double pad) {
  std::vector<double> vx(x, x + n), vy(y, y + n), vp(phi, phi + n), va(A, A + n), vr(R, R + n);
  return new SpherocylBox(L, n, vx.data(), vy.data(), vp.data(), va.data(), vr.data(), pad);
}

void ref_box_free(void* h) { delete static_cast<SpherocylBox*>(h); }

void ref_set_params(void* h, double gamma, double gamma_total, double strain_rate,
                    double resize_rate, double step) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  b->m_dGamma = gamma;
  b->m_dGammaTotal = gamma_total;
  b->m_dStrainRate = strain_rate;
  b->m_dResizeRate = resize_rate;
  b->m_dStep = step;
}

void ref_set_save_steps(void* h, double pos_step, double se_step) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  b->m_dPosSaveStep = pos_step;
  b->m_dSESaveStep = se_step;
}

void ref_calc_se(void* h) { static_cast<SpherocylBox*>(h)->calc_se(); }
void ref_calc_forces(void* h) { static_cast<SpherocylBox*>(h)->calc_forces(); }
void ref_find_neighbors(void* h) { static_cast<SpherocylBox*>(h)->find_neighbors(); }
void ref_set_back_gamma(void* h) { static_cast<SpherocylBox*>(h)->set_back_gamma(); }
void ref_strain_step(void* h) { static_cast<SpherocylBox*>(h)->strain_step(); }
void ref_resize_step(void* h, int shrink) { static_cast<SpherocylBox*>(h)->resize_step(shrink != 0); }
void ref_gd_step(void* h) { static_cast<SpherocylBox*>(h)->gradient_relax_step(); }

// Per-particle state in the reference's own (unsorted) particle order.
void ref_get_particles(void* h, double* x, double* y, double* phi, double* R, double* A,
                       double* fx, double* fy, double* tau, double* I, double* RC) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  for (int i = 0; i < b->m_nParticles; ++i) {
    const Spherocyl& s = b->m_psParticles[i];
    if (x) x[i] = s.m_dX;
    if (y) y[i] = s.m_dY;
    if (phi) phi[i] = s.m_dPhi;
    if (R) R[i] = s.m_dR;
    if (A) A[i] = s.m_dA;
    if (fx) fx[i] = s.m_dFx;
    if (fy) fy[i] = s.m_dFy;
    if (tau) tau[i] = s.m_dTau;
    if (I) I[i] = s.m_dI;
    if (RC) RC[i] = s.m_dRC;
  }
}

// out[0..15]: Lx Ly gamma gamma_total strain_rate resize_rate step packing
//             padding E Pxx Pyy Pxy Pyx rows cols
void ref_get_scalars(void* h, double* out) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  double v[16] = {b->m_dLx, b->m_dLy, b->m_dGamma, b->m_dGammaTotal, b->m_dStrainRate,
                  b->m_dResizeRate, b->m_dStep, b->m_dPacking, b->m_dPadding, b->m_dEnergy,
                  b->m_dPxx, b->m_dPyy, b->m_dPxy, b->m_dPyx, (double)b->m_nCellRows,
                  (double)b->m_nCellCols};
  std::memcpy(out, v, sizeof(v));
}

// Half neighbour list as (p, q) pairs, duplicates kept as the reference stores them.
// Returns the total number of entries (may exceed cap; only cap are written).
int ref_get_neighbors(void* h, int* p_out, int* q_out, int cap) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  int k = 0;
  for (int p = 0; p < b->m_nParticles; ++p) {
    for (size_t j = 0; j < b->m_pvNeighbors[p].size(); ++j) {
      if (k < cap) {
        p_out[k] = p;
        q_out[k] = b->m_pvNeighbors[p][j];
      }
      ++k;
    }
  }
  return k;
}

// The inner loop of run_strain without file I/O (src/cpp/strain_dynamics.cpp:285-296):
// strain_step, then calc_se every se_every sub-steps (recording E Pxx Pyy Pxy) or
// calc_forces otherwise.  se_out receives 4 doubles per recorded sub-step.
int ref_run_steps(void* h, int nsteps, int se_every, double* se_out) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  int rec = 0;
  for (int s = 1; s <= nsteps; ++s) {
    b->strain_step();
    if (se_every > 0 && s % se_every == 0) {
      b->calc_se();
      if (se_out) {
        se_out[4 * rec + 0] = b->m_dEnergy;
        se_out[4 * rec + 1] = b->m_dPxx;
        se_out[4 * rec + 2] = b->m_dPyy;
        se_out[4 * rec + 3] = b->m_dPxy;
      }
      ++rec;
    } else {
      b->calc_forces();
    }
  }
  return rec;
}

// Full protocol entry points with the reference's own file output.
unsigned long ref_run_strain(void* h, double run_length, unsigned long ntime, const char* outdir,
                             const char* sefile) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  b->set_output_directory(outdir);
  b->set_se_file(sefile);
  return b->run_strain(run_length, ntime);
}

unsigned long ref_resize_box(void* h, double final_packing, unsigned long ntime, const char* outdir,
                             const char* sefile) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  b->set_output_directory(outdir);
  b->set_se_file(sefile);
  return b->resize_box(final_packing, ntime);
}

void ref_gd_minimize(void* h, int nmax, double emin, const char* outdir, const char* sefile) {
  SpherocylBox* b = static_cast<SpherocylBox*>(h);
  b->set_output_directory(outdir);
  b->set_se_file(sefile);
  b->gradient_descent_minimize(nmax, emin);
}

}  // extern "C"
