// Device-side physics shared by the tile and list force kernels: the pair
// force, minimum image, wrapping, cell binning and block reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "sc_internal.h"

namespace sc {

// ---------------------------------------------------------------------------
// pair math
// ---------------------------------------------------------------------------
template <typename T>
struct PairOut {
  T fx, fy, tp, tq, s, t, dx, dy, dist, dV;
};

// Segment-segment closest approach and harmonic contact force, following
// calc_force_pair (src/cpp/strain_dynamics.cpp:38-64) term by term.  Returns
// false when |D| >= sigma (or D is NaN).  FP64: no contraction, so the rounding
// matches the reference's x86-64 build.
__device__ __forceinline__ bool pair_contact_f64(double ddx, double ddy, double nxA, double nyA,
                                                 double nxB, double nyB, double Ap, double Aq,
                                                 double sig, PairOut<double>& o) {
#pragma clang fp contract(off)
  double a = Ap * Ap;
  double b = -(nxA * nxB + nyA * nyB);
  double c = Aq * Aq;
  double d = nxA * ddx + nyA * ddy;
  double e = -nxB * ddx - nyB * ddy;
  double det = a * c - b * b;
  double t = fmin(fmax((b * d - a * e) / det, -1.0), 1.0);
  double s = -(b * t + d) / a;
  double sabs = fabs(s);
  s = fmin(fmax(s, -1.0), 1.0);
  if (sabs > 1.0) t = fmin(fmax(-(b * s + e) / c, -1.0), 1.0);
  double Dx = ddx + s * nxA - t * nxB;
  double Dy = ddy + s * nyA - t * nyB;
  double d2 = Dx * Dx + Dy * Dy;
  if (!(d2 < sig * sig)) return false;
  double dist = sqrt(d2);
  double dV = (1.0 - dist / sig) / sig;
  double fx = Dx * dV / dist;
  double fy = Dy * dV / dist;
  o.fx = fx;
  o.fy = fy;
  o.tp = s * nxA * fy - s * nyA * fx;
  o.tq = t * nyB * fx - t * nxB * fy;
  o.s = s;
  o.t = t;
  o.dx = Dx;
  o.dy = Dy;
  o.dist = dist;
  o.dV = dV;
  return true;
}

// The same closest-approach force in FP32 (mixed mode), contraction allowed.
__device__ __forceinline__ bool pair_contact_f32(float ddx, float ddy, float nxA, float nyA,
                                                 float nxB, float nyB, float Ap, float Aq,
                                                 float sig, PairOut<float>& o) {
#pragma clang fp contract(fast)
  float a = Ap * Ap;
  float b = -(nxA * nxB + nyA * nyB);
  float c = Aq * Aq;
  float d = nxA * ddx + nyA * ddy;
  float e = -nxB * ddx - nyB * ddy;
  float det = a * c - b * b;
  float t = fminf(fmaxf((b * d - a * e) / det, -1.0f), 1.0f);
  float s = -(b * t + d) / a;
  float sabs = fabsf(s);
  s = fminf(fmaxf(s, -1.0f), 1.0f);
  if (sabs > 1.0f) t = fminf(fmaxf(-(b * s + e) / c, -1.0f), 1.0f);
  float Dx = ddx + s * nxA - t * nxB;
  float Dy = ddy + s * nyA - t * nyB;
  float d2 = Dx * Dx + Dy * Dy;
  if (!(d2 < sig * sig)) return false;
  float dist = sqrtf(d2);
  float dV = (1.0f - dist / sig) / sig;
  float fx = Dx * dV / dist;
  float fy = Dy * dV / dist;
  o.fx = fx;
  o.fy = fy;
  o.tp = s * nxA * fy - s * nyA * fx;
  o.tq = t * nyB * fx - t * nxB * fy;
  o.s = s;
  o.t = t;
  o.dx = Dx;
  o.dy = Dy;
  o.dist = dist;
  o.dV = dV;
  return true;
}

// Minimum image in sheared coordinates (src/cpp/strain_dynamics.cpp:28-29).
__device__ __forceinline__ double min_image(double d, double L) {
#pragma clang fp contract(off)
  return d + L * (double)((d < -0.5 * L) - (d > 0.5 * L));
}

// Wrap rule of find_cells (src/cpp/spherocyl_box.cpp:279-286); bounded so a
// runaway coordinate cannot spin.  Returns false for a non-finite coordinate.
__device__ __forceinline__ bool wrap_coord(double& v, double L) {
  if (!isfinite(v)) return false;
  int guard = 0;
  while (v > L && guard < 8) { v -= L; ++guard; }
  while (v < 0.0 && guard < 16) { v += L; ++guard; }
  if (guard >= 8) v = v - L * floor(v / L);
  return true;
}

__device__ __forceinline__ void wrap_angle(double& p) {
  if (!isfinite(p)) return;
  int guard = 0;
  while (p > kRefPi && guard < 8) { p -= 2 * kRefPi; ++guard; }
  while (p < -kRefPi && guard < 16) { p += 2 * kRefPi; ++guard; }
}

// Bin one particle into bucket buffer (cnt, slot); the cell rule is find_cells'
// (src/cpp/spherocyl_box.cpp:293-295), clamped so x == L stays in range.
__device__ __forceinline__ void bin_into(int g, double xw, double yw, double Lx, double Ly,
                                         const PackInfo& P, int* cnt, int* slot, int* err) {
  double w = Lx / P.cols, h = Ly / P.rows;
  int col = (int)(xw / w);
  int row = (int)(yw / h);
  col = min(max(col, 0), P.cols - 1);
  row = min(max(row, 0), P.rows - 1);
  int c = P.coff + row * P.cols + col;
  int k = atomicAdd(&cnt[c], 1);
  if (k < kCap) slot[c * kCap + k] = g;
  else atomicOr(err, kErrCellOverflow);
}

// isolated_rotation, src/cpp/strain_dynamics.cpp:174-177
__device__ __forceinline__ double iso_rot(double RC, double phi) {
#pragma clang fp contract(off)
  return 0.5 * (1 - RC * cos(2 * phi));
}

// ---------------------------------------------------------------------------
// block reduction helpers (fixed order -> deterministic)
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* s_red, int nthr) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double x = v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    v[i] = x;
  }
  const int w = tid >> 6, nw = nthr >> 6;
  if ((tid & 63) == 0)
#pragma unroll
    for (int i = 0; i < NV; ++i) s_red[w * NV + i] = v[i];
  __syncthreads();
  if (tid < NV) {
    double x = 0;
    for (int k = 0; k < nw; ++k) x += s_red[k * NV + tid];
    s_red[tid] = x;  // each slot tid is read (k = 0) before this write by the same thread
  }
  __syncthreads();
}

}  // namespace sc
