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

// Crossed segments (their lines meet inside both): the closest distance is 0
// and stays 0 under small displacements, and D is rounding noise.  The overlap
// energy is continued to negative distance d = -delta, delta the minimal
// translation that uncrosses them: the smallest distance of an endpoint from
// the other segment's line.  That endpoint's segment is pushed back across
// the line; the force is (1 + delta/sigma)/sigma along the push, at the
// endpoint and at its projection on the other segment -- the ordinary contact
// force of the configuration it uncrosses into, so force and energy are
// continuous where the rods uncross.  Fills o as a contact with dist = -delta.
template <typename T>
__device__ __forceinline__ void crossed_contact(T ddx, T ddy, T nxA, T nyA, T nxB, T nyB, T sig,
                                                PairOut<T>& o) {
#pragma clang fp contract(off)
  const T la2 = nxA * nxA + nyA * nyA, lb2 = nxB * nxB + nyB * nyB;
  const T la = sqrt(la2), lb = sqrt(lb2);
  // unit normals of the segments
  const T mAx = -nyA / la, mAy = nxA / la, mBx = -nyB / lb, mBy = nxB / lb;
  // A's endpoints (relative to B's centre) against B's line, and B's against A's
  const T hp = (ddx + nxA) * mBx + (ddy + nyA) * mBy;
  const T hm = (ddx - nxA) * mBx + (ddy - nyA) * mBy;
  const T gp = (-ddx + nxB) * mAx + (-ddy + nyB) * mAy;
  const T gm = (-ddx - nxB) * mAx + (-ddy - nyB) * mAy;
  const T dA = fmin(fabs(hp), fabs(hm)), dB = fmin(fabs(gp), fabs(gm));
  T s, t, ux, uy, delta;
  if (dA <= dB) {  // move A: its nearer endpoint back across B's line
    s = fabs(hp) <= fabs(hm) ? T(1) : T(-1);
    const T h = s > 0 ? hp : hm;
    ux = h > 0 ? -mBx : mBx;
    uy = h > 0 ? -mBy : mBy;
    t = fmin(fmax(((ddx + s * nxA) * nxB + (ddy + s * nyA) * nyB) / lb2, T(-1)), T(1));
    delta = dA;
  } else {  // move B: its nearer endpoint back across A's line (A gets -F_B)
    t = fabs(gp) <= fabs(gm) ? T(1) : T(-1);
    const T gg = t > 0 ? gp : gm;
    ux = gg > 0 ? mAx : -mAx;
    uy = gg > 0 ? mAy : -mAy;
    s = fmin(fmax(((-ddx + t * nxB) * nxA + (-ddy + t * nyB) * nyA) / la2, T(-1)), T(1));
    delta = dB;
  }
  const T dV = (T(1) + delta / sig) / sig;
  const T fx = ux * dV, fy = uy * dV;
  o.fx = fx;
  o.fy = fy;
  o.tp = s * nxA * fy - s * nyA * fx;
  o.tq = t * nyB * fx - t * nxB * fy;
  o.s = s;
  o.t = t;
  o.dist = -delta;
  o.dx = -delta * ux;
  o.dy = -delta * uy;
  o.dV = dV;
}

// Whether two segments intersect: each one's endpoints lie strictly on
// opposite sides of the other's line (cross products, no normalisation).  The
// closest-approach test of crossing (|tu| < 1 and |s| < 1) also fires on noise
// for nearly parallel segments, whose det = a c - b^2 loses its digits to
// cancellation (in FP32 from angles of ~1e-3 down); a crossing is taken only
// when this test confirms it.  For such pairs the endpoints sit a contact
// distance from the other line, far above the rounding of these products.
template <typename T>
__device__ __forceinline__ bool segments_cross(T ddx, T ddy, T nxA, T nyA, T nxB, T nyB) {
#pragma clang fp contract(off)
  const T hp = nxB * (ddy + nyA) - nyB * (ddx + nxA);
  const T hm = nxB * (ddy - nyA) - nyB * (ddx - nxA);
  const T gp = nxA * (nyB - ddy) - nyA * (nxB - ddx);
  const T gm = nxA * (-nyB - ddy) - nyA * (-nxB - ddx);
  return ((hp > T(0) && hm < T(0)) || (hp < T(0) && hm > T(0))) &&
         ((gp > T(0) && gm < T(0)) || (gp < T(0) && gm > T(0)));
}

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
  double tu = (b * d - a * e) / det;
  double t = fmin(fmax(tu, -1.0), 1.0);
  double s = -(b * t + d) / a;
  double sabs = fabs(s);
  s = fmin(fmax(s, -1.0), 1.0);
  if (sabs > 1.0) t = fmin(fmax(-(b * s + e) / c, -1.0), 1.0);
  double Dx = ddx + s * nxA - t * nxB;
  double Dy = ddy + s * nyA - t * nyB;
  double d2 = Dx * Dx + Dy * Dy;
  if (!(d2 < sig * sig)) return false;
  o.s = s;
  o.t = t;
  o.dx = Dx;
  o.dy = Dy;
  if (det > 0.0 && fabs(tu) < 1.0 && sabs < 1.0 &&
      segments_cross(ddx, ddy, nxA, nyA, nxB, nyB)) {
    crossed_contact(ddx, ddy, nxA, nyA, nxB, nyB, sig, o);  // crossed segments
    return true;
  }
  if (d2 == 0.0) {  // touching exactly (ends, collinear overlap): no direction
    o.fx = o.fy = o.tp = o.tq = 0.0;
    o.dist = 0.0;
    o.dV = 1.0 / sig;
    return true;
  }
  double dist = sqrt(d2);
  double dV = (1.0 - dist / sig) / sig;
  double fx = Dx * dV / dist;
  double fy = Dy * dV / dist;
  o.fx = fx;
  o.fy = fy;
  o.tp = s * nxA * fy - s * nyA * fx;
  o.tq = t * nyB * fx - t * nxB * fy;
  o.dist = dist;
  o.dV = dV;
  return true;
}

// ---------------------------------------------------------------------------
// two-point contact model
// ---------------------------------------------------------------------------
// Force fields of one contact point from its D vector (harmonic, as above).
template <typename T>
__device__ __forceinline__ bool contact_point(T Dx, T Dy, T s, T t, T nxA, T nyA, T nxB, T nyB,
                                              T sig, bool crossed, T ddx, T ddy, PairOut<T>& o) {
#pragma clang fp contract(off)
  T d2 = Dx * Dx + Dy * Dy;
  if (crossed) {  // as in pair_contact_f64
    crossed_contact(ddx, ddy, nxA, nyA, nxB, nyB, sig, o);
    return true;
  }
  if (!(d2 < sig * sig)) return false;
  o.s = s;
  o.t = t;
  o.dx = Dx;
  o.dy = Dy;
  if (d2 == T(0)) {
    o.fx = o.fy = o.tp = o.tq = T(0);
    o.dist = T(0);
    o.dV = T(1) / sig;
    return true;
  }
  T dist = sqrt(d2);
  T dV = (T(1) - dist / sig) / sig;
  T fx = Dx * dV / dist;
  T fy = Dy * dV / dist;
  o.fx = fx;
  o.fy = fy;
  o.tp = s * nxA * fy - s * nyA * fx;
  o.tq = t * nyB * fx - t * nxB * fy;
  o.dist = dist;
  o.dV = dV;
  return true;
}

// Mixed precision (FP32 pair forces).  The closest-approach parameters s, t
// and the crossing decision stay FP64, in pair_contact_f64's order: they come
// from differences of nearly equal products (det = a c - b^2, b d - a e), which
// in FP32 are noise for nearly parallel rods -- a noise tu in (-1, 1) would
// put the contact anywhere along the overlap, or call the pair crossed.  The
// contact vector D is formed in FP64 as well; its length, the overlap force
// and the torques are FP32.
__device__ __forceinline__ bool pair_contact_mixed(double ddx, double ddy, double nxA, double nyA,
                                                   double nxB, double nyB, double Ap, double Aq,
                                                   double sig, PairOut<float>& o) {
#pragma clang fp contract(off)
  double a = Ap * Ap;
  double b = -(nxA * nxB + nyA * nyB);
  double c = Aq * Aq;
  double d = nxA * ddx + nyA * ddy;
  double e = -nxB * ddx - nyB * ddy;
  double det = a * c - b * b;
  double tu = (b * d - a * e) / det;
  double t = fmin(fmax(tu, -1.0), 1.0);
  double s = -(b * t + d) / a;
  double sabs = fabs(s);
  s = fmin(fmax(s, -1.0), 1.0);
  if (sabs > 1.0) t = fmin(fmax(-(b * s + e) / c, -1.0), 1.0);
  const bool crossed = det > 0.0 && fabs(tu) < 1.0 && sabs < 1.0 &&
                       segments_cross(ddx, ddy, nxA, nyA, nxB, nyB);
  return contact_point<float>((float)(ddx + s * nxA - t * nxB), (float)(ddy + s * nyA - t * nyB),
                              (float)s, (float)t, (float)nxA, (float)nyA, (float)nxB, (float)nyB,
                              (float)sig, crossed, (float)ddx, (float)ddy, o);
}

template <typename T>
__device__ __forceinline__ T clamp1(T v) {
  return fmin(fmax(v, T(-1)), T(1));
}

// Two-point contact model (src/cuda/2point_dynamics.cu:391-546).  Besides the
// closest pair of points (s1, t1), a second pair (s2, t2) -- for nearly parallel
// rods the other end of their overlap -- carries its own harmonic contact when
// it overlaps as well, so two overlapping, nearly aligned rods have no energy
// kink at alignment.  The second pair starts at the opposite end of whichever
// segment fixed the first (s2 = -s1 or t2 = -t1) projected onto the other
// segment; a third candidate from the other segment's matching end replaces it
// when it is closer.  o0 / h0: the closest-point contact; o1 / h1: the second.
// The candidate parameters, every choice between them and the contact vectors
// are FP64; T is the type of each contact's force evaluation (FP32 in mixed
// mode, see pair_contact_mixed).
template <typename T>
__device__ __forceinline__ void pair_contacts_2p(double ddx, double ddy, double nxA, double nyA,
                                                 double nxB, double nyB, double Ap, double Aq,
                                                 double sig, PairOut<T>& o0, bool& h0,
                                                 PairOut<T>& o1, bool& h1) {
#pragma clang fp contract(off)
  const double a = Ap * Ap;
  const double b = -(nxA * nxB + nyA * nyB);
  const double c = Aq * Aq;
  const double d = nxA * ddx + nyA * ddy;
  const double e = -nxB * ddx - nyB * ddy;
  const double det = a * c - b * b;
  double s1, t1, s2, t2, s3, t3;
  bool crossed = false;
  if (det <= 0.0) {  // parallel: the two ends of the overlap
    s1 = clamp1(-(b + d) / a);
    s2 = clamp1(-(d - b) / a);
    if (fabs(s1) == 1.0) t1 = clamp1(-(b * s1 + e) / c);
    else t1 = b < 0.0 ? clamp1(-(b + e) / c) : clamp1(-(e - b) / c);
    if (fabs(s2) == 1.0) t2 = clamp1(-(b * s2 + e) / c);
    else t2 = b < 0.0 ? clamp1(-(e - b) / c) : clamp1(-(b + e) / c);
    s3 = s2;
    t3 = t2;
  } else {
    const double tu = (b * d - a * e) / det;
    t1 = clamp1(tu);
    s1 = -(b * t1 + d) / a;
    double sarg = fabs(s1);
    s1 = clamp1(s1);
    crossed = fabs(tu) < 1.0 && sarg < 1.0 && segments_cross(ddx, ddy, nxA, nyA, nxB, nyB);
    if (sarg >= 1.0) {
      t1 = clamp1(-(b * s1 + e) / c);
      s2 = -s1;
      t2 = -(b * s2 + e) / c;
      double targ = fabs(t2);
      t2 = clamp1(t2);
      if (targ > 1.0) {
        s2 = clamp1(-(b * t2 + d) / a);
        s3 = s2;
        t3 = t2;
      } else {
        t3 = b < 0.0 ? -s1 : s1;
        s3 = -(b * t3 + d) / a;
        sarg = fabs(s3);
        s3 = clamp1(s3);
        if (sarg >= 1.0) t3 = clamp1(-(b * s3 + e) / c);
      }
    } else {
      t2 = -t1;
      s2 = -(b * t2 + d) / a;
      sarg = fabs(s2);
      s2 = clamp1(s2);
      if (sarg > 1.0) {
        t2 = clamp1(-(b * s2 + e) / c);
        s3 = s2;
        t3 = t2;
      } else {
        s3 = b < 0.0 ? -t1 : t1;
        t3 = -(b * s3 + e) / c;
        double targ = fabs(t3);
        t3 = clamp1(t3);
        if (targ >= 1.0) s3 = clamp1(-(b * t3 + d) / a);
      }
    }
  }
  h0 = contact_point<T>(T(ddx + s1 * nxA - t1 * nxB), T(ddy + s1 * nyA - t1 * nyB), T(s1), T(t1),
                        T(nxA), T(nyA), T(nxB), T(nyB), T(sig), crossed, T(ddx), T(ddy), o0);
  h1 = false;
  if (s1 != s2 && t1 != t2) {
    double Dx2 = ddx + s2 * nxA - t2 * nxB;
    double Dy2 = ddy + s2 * nyA - t2 * nyB;
    if (s3 != s1 && s3 != s2) {
      const double Dx3 = ddx + s3 * nxA - t3 * nxB;
      const double Dy3 = ddy + s3 * nyA - t3 * nyB;
      if (Dx3 * Dx3 + Dy3 * Dy3 < Dx2 * Dx2 + Dy2 * Dy2) {
        s2 = s3;
        t2 = t3;
        Dx2 = Dx3;
        Dy2 = Dy3;
      }
    }
    h1 = contact_point<T>(T(Dx2), T(Dy2), T(s2), T(t2), T(nxA), T(nyA), T(nxB), T(nyB), T(sig),
                          false, T(ddx), T(ddy), o1);
  }
}

template <typename T>
__device__ __forceinline__ void widen(const PairOut<T>& a, PairOut<double>& o) {
  o.fx = a.fx; o.fy = a.fy; o.tp = a.tp; o.tq = a.tq; o.s = a.s; o.t = a.t;
  o.dx = a.dx; o.dy = a.dy; o.dist = a.dist; o.dV = a.dV;
}

// Contacts of one pair in the orientation given (A = lower id): the
// closest-point model of the CPU reference, or (TWO_POINT) the two-point model;
// FP64 or mixed (FP32 force evaluation) pair math, FP64 results.  o0 / h0:
// closest-point contact; o1 / h1: the second contact of the two-point model.
template <bool TWO_POINT>
__device__ __forceinline__ void pair_contacts(double ddx, double ddy, double nxA, double nyA,
                                              double nxB, double nyB, double Ap, double Aq,
                                              double sig, bool fp64, PairOut<double>& o0,
                                              bool& h0, PairOut<double>& o1, bool& h1) {
  if (!TWO_POINT) {
    h1 = false;
    if (fp64) {
      h0 = pair_contact_f64(ddx, ddy, nxA, nyA, nxB, nyB, Ap, Aq, sig, o0);
    } else {
      PairOut<float> f0;
      h0 = pair_contact_mixed(ddx, ddy, nxA, nyA, nxB, nyB, Ap, Aq, sig, f0);
      if (h0) widen(f0, o0);
    }
    return;
  }
  if (fp64) {
    pair_contacts_2p<double>(ddx, ddy, nxA, nyA, nxB, nyB, Ap, Aq, sig, o0, h0, o1, h1);
  } else {
    PairOut<float> f0, f1;
    pair_contacts_2p<float>(ddx, ddy, nxA, nyA, nxB, nyB, Ap, Aq, sig, f0, h0, f1, h1);
    if (h0) widen(f0, o0);
    if (h1) widen(f1, o1);
  }
}

// Stress / energy sums of one contact point of pair (A, B) (A = lower id):
// the reference's formulas with their two slips (src/cpp/spherocyl_box.cpp:
// 483-495) and the corrected virial, unnormalised (k_reduce divides).
__device__ __forceinline__ void add_contact_sums(const PairOut<double>& r, double nxA, double nyA,
                                                 double nxB, double nyB, double sig, double* acc) {
#pragma clang fp contract(off)
  double cxa = r.s * nxA - 0.5 * r.dx;
  double cya = r.s * nyA - 0.5 * r.dy;
  double cxb = r.t * nxB + 0.5 * r.dx;
  double cyb_ref = r.t * nxB + 0.5 * r.dy;
  double cyb = r.t * nyB + 0.5 * r.dy;
  acc[kSeE] += r.dV * sig * (1.0 - r.dist / sig);
  acc[kSePxx] += r.fx * cxa - r.fx * cxb - r.fy * cxb;
  acc[kSePyy] += r.fy * cya - r.fy * cyb_ref;
  acc[kSePxy] += r.fy * cxa;
  acc[kSeCPxx] += r.fx * (cxa - cxb);
  acc[kSeCPyy] += r.fy * (cya - cyb);
  acc[kSeCPxy] += r.fy * (cxa - cxb);
  acc[kSeCPyx] += r.fx * (cya - cyb);
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

// Column and row of a wrapped position: find_cells' floor(x / w)
// (src/cpp/spherocyl_box.cpp:293-295) as floor(x cols / Lx), one division
// instead of two, clamped so x == L stays in range.  Every kernel that bins a
// particle or looks up its own cell uses this one rule, so a particle's cell
// is the same wherever it is computed.
__device__ __forceinline__ void cell_col_row(double xw, double yw, double Lx, double Ly,
                                             const PackInfo& P, int& col, int& row) {
  col = min(max((int)(xw * ((double)P.cols / Lx)), 0), P.cols - 1);
  row = min(max((int)(yw * ((double)P.rows / Ly)), 0), P.rows - 1);
}

__device__ __forceinline__ int cell_of(double xw, double yw, double Lx, double Ly,
                                      const PackInfo& P) {
  int col, row;
  cell_col_row(xw, yw, Lx, Ly, P, col, row);
  return P.coff + row * P.cols + col;
}

// Bin one particle into bucket buffer (cnt, slot).
__device__ __forceinline__ void bin_into(int g, double xw, double yw, double Lx, double Ly,
                                         const PackInfo& P, int* cnt, int* slot, int* err) {
  const int c = cell_of(xw, yw, Lx, Ly, P);
  int k = atomicAdd(&cnt[c], 1);
  if (k < P.cap) slot[(size_t)c * P.cap + k] = g;
  else atomicOr(err, kErrCellOverflow);
}

// Wavefront-aggregated bucket fill.  Every lane of the wavefront calls it
// (uniform control flow); lane l with c >= 0 puts particle g into bucket c.
// Particles are handled by lanes STRIDE apart (lane 0 of each particle's
// lanes), and storage is sorted by cell, so neighbouring particles often share
// a cell: each run of such lanes with one cell takes its slots with a single
// atomic, issued by the run's first lane, and the lanes of the run take
// consecutive slots from the returned base.  All runs' atomics are issued
// together.  Slot order within a bucket carries no meaning (rows and re-sorts
// order members by particle id).  It pays in the minimiser update, whose
// wavefronts do little besides the binning (about 2.5 particles per cell);
// the Heun passes keep one atomic per particle, which measured no slower there.
template <int STRIDE>
__device__ __forceinline__ void bin_wave(int c, int g, int cap, int* cnt, int* slot, int* err) {
  const int lane = __lane_id();
  const bool act = c >= 0;
  const int cp = __shfl_up(c, STRIDE);  // lanes below STRIDE read their own value
  const bool head = act && (lane < STRIDE || cp != c);
  const unsigned long long actm = __ballot(act), heads = __ballot(head);
  const unsigned long long below = (1ull << lane) - 1ull;  // lanes < lane
  const unsigned long long upto = (2ull << lane) - 1ull;   // lanes <= lane (all for lane 63)
  // first lane of this lane's run (the highest head at or below it)
  const unsigned long long hm = heads & upto;
  const int h = act && hm ? 63 - __clzll((long long)hm) : lane;
  int base = 0;
  if (head) {
    const unsigned long long after = heads & ~upto;  // heads above this lane
    const int end = after ? __ffsll((long long)after) - 1 : 64;
    const unsigned long long span = (end == 64 ? ~0ull : ((1ull << end) - 1ull)) & ~below;
    base = atomicAdd(&cnt[c], __popcll(actm & span));
  }
  base = __shfl(base, h);
  if (act) {
    const int k = base + __popcll(actm & below & ~((1ull << h) - 1ull));
    if (k < cap) slot[(size_t)c * cap + k] = g;
    else atomicOr(err, kErrCellOverflow);
  }
}

// isolated_rotation, src/cpp/strain_dynamics.cpp:174-177: 0.5 (1 - RC cos 2 phi),
// with cos 2 phi taken from the segment vector (nx, ny) = A (cos phi, sin phi)
// the state carries, (nx^2 - ny^2) / A^2, instead of an FP64 cosine (a disc,
// A = 0, has RC = 0).  It differs from cos(2 phi) by rounding only.
__device__ __forceinline__ double iso_rot_seg(double RC, double nx, double ny, double A) {
#pragma clang fp contract(off)
  const double c2 = A > 0.0 ? (nx * nx - ny * ny) / (A * A) : 1.0;
  return 0.5 * (1 - RC * c2);
}

// Force of partner q on particle i (i's data in registers), accumulated into
// (Fx, Fy, Tq) and, for the lower-id side of the pair, into the stress/energy
// sums.  The pair is evaluated in the reference's orientation (lower particle
// id first, src/cpp/spherocyl_box.cpp:514-520) so both sides agree bitwise.
template <bool SUMS, bool TWO_POINT>
__device__ __forceinline__ void pair_into(double xi, double yi, double nxi, double nyi, double Ri,
                                          double Ai, int idi, double xq, double yq, double nxq,
                                          double nyq, double Rq, double Aq, int idq, double Lx,
                                          double Ly, double gam, bool fp64, bool do_se,
                                          double& Fx, double& Fy, double& Tq, double* acc) {
#pragma clang fp contract(off)
  const bool iP = idi < idq;
  double ddx = min_image(iP ? xi - xq : xq - xi, Lx);
  double ddy = min_image(iP ? yi - yq : yq - yi, Ly);
  ddx += gam * ddy;
  const double sig = Ri + Rq;
  const double nxA = iP ? nxi : nxq, nyA = iP ? nyi : nyq;
  const double nxB = iP ? nxq : nxi, nyB = iP ? nyq : nyi;
  const double Ap = iP ? Ai : Aq, Aqq = iP ? Aq : Ai;
  PairOut<double> r0, r1;
  bool h0, h1;
  pair_contacts<TWO_POINT>(ddx, ddy, nxA, nyA, nxB, nyB, Ap, Aqq, sig, fp64, r0, h0, r1, h1);
  if (h0) {
    if (iP) {
      Fx += r0.fx;
      Fy += r0.fy;
      Tq += r0.tp;
      if (SUMS && do_se) {
        add_contact_sums(r0, nxA, nyA, nxB, nyB, sig, acc);
        acc[kSeNc] += 1.0;
      }
    } else {
      Fx -= r0.fx;
      Fy -= r0.fy;
      Tq += r0.tq;
    }
  }
  if (TWO_POINT && h1) {
    if (iP) {
      Fx += r1.fx;
      Fy += r1.fy;
      Tq += r1.tp;
      if (SUMS && do_se) add_contact_sums(r1, nxA, nyA, nxB, nyB, sig, acc);
    } else {
      Fx -= r1.fx;
      Fy -= r1.fy;
      Tq += r1.tq;
    }
  }
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
