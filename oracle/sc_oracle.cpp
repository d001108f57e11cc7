// ORACLE -- test infrastructure only.
//
// A CPU restatement of the spherocylinder dynamics of tmarschall/spherocyls
// (src/cpp), written from the equations and operation order of that code so
// that it agrees with the compiled reference to rounding.  The HIP path is
// tested against it, and it is the timed CPU baseline ("port"); nothing on the
// product path may link or call it.
//
// Pinning (what each part is checked against, see tests/):
//   pair force / torque / E / stress .. golden vectors from the compiled
//       reference (tests/golden), two-body closed forms, sum(F)=0,
//       F = -N dE/dx by finite differences
//   cells / neighbour list ............ golden neighbour sets, O(N^2) brute force
//   Heun / resize / gradient steps .... golden one-step and trajectory states
//   FIRE .............................. no reference counterpart: pinned by its
//       defining properties (residual reaches tolerance, energy never rises on a
//       step that keeps the velocity, reduces to gradient descent for alpha=1)
#include "sc_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

// The reference uses a truncated pi everywhere (src/cpp/spherocyl_box.cpp:12).
constexpr double kPi = 3.14159265358979;

struct Cells {
  int rows = 1, cols = 1;
  double w = 0, h = 0;
  std::vector<int> adj;                 // 8 per cell
  std::vector<std::vector<int>> members;
};

}  // namespace

struct orc_box {
  int n = 0;
  double Lx = 0, Ly = 0, gamma = 0, gamma_total = 0, strain_rate = 0, resize_rate = 0;
  double dt = 0.05, pad = 0, packing = 0, Rmax = 0, Amax = 0;
  std::vector<double> x, y, phi, R, A, I, RC, fx, fy, tau, xm, ym;
  std::vector<double> tx, ty, tphi, tfx, tfy, ttau;  // predictor copies
  std::vector<int> cell;
  Cells cg;
  std::vector<std::vector<int>> nbr;  // half lists (q > p), duplicates kept
  double E = 0, Pxx = 0, Pyy = 0, Pxy = 0;           // reference formulas
  double cPxx = 0, cPyy = 0, cPxy = 0, cPyx = 0;     // corrected virial
  int ncontacts = 0;
  int contact_model = 0;  // 0: closest point only (src/cpp), 1: two-point model
};

namespace {

// set_particle's rotational constants, spherocyl_box.cpp:198-202
void rot_consts(double R, double A, double* I, double* RC) {
  double al = A / R;
  double C = (3 * kPi + 24 * al + 6 * kPi * al * al + 8 * al * al * al);
  double area = R * (kPi * R + 4 * A);
  *I = R * R * R * R * C / (6 * area);
  *RC = (8 * al + 6 * kPi * al * al + 8 * al * al * al) / C;
}

// Wrap rule of set_particle / find_cells (spherocyl_box.cpp:204-215, 279-290):
// strict comparisons, so a coordinate equal to L stays at L.
inline void wrap_pos(double& v, double L) {
  while (v > L) v -= L;
  while (v < 0.0) v += L;
}
inline void wrap_angle(double& p) {
  while (p > kPi) p -= 2 * kPi;
  while (p < -kPi) p += 2 * kPi;
}

// configure_cells, spherocyl_box.cpp:223-268.  For Lx != Ly (which the CPU
// constructors never produce) the column width grows to cover the remap strain
// Lx / (2 Ly): 2 sqrt(1 + (Lx / 2 Ly)^2) (Rmax + Amax) + pad when above the
// reference's 2.24 (Rmax + Amax) + pad.
void configure_cells(orc_box* b) {
  const double r = b->Lx / b->Ly;
  double wmin = std::max(2.24, 2.0 * std::sqrt(1.0 + 0.25 * r * r)) * (b->Rmax + b->Amax) + b->pad;
  double hmin = 2 * (b->Rmax + b->Amax) + b->pad;
  Cells& c = b->cg;
  c.rows = std::max(static_cast<int>(b->Ly / hmin), 1);
  c.cols = std::max(static_cast<int>(b->Lx / wmin), 1);
  c.w = b->Lx / c.cols;
  c.h = b->Ly / c.rows;
  if (b->pad == 0) b->pad = std::max(c.w - wmin, c.h - hmin);
  int nc = c.rows * c.cols;
  c.members.assign(nc, {});
  c.adj.assign(8 * nc, 0);
  for (int k = 0; k < nc; ++k) {
    int r = k / c.cols, col = k % c.cols;
    int cr = (col + 1) % c.cols, cl = (c.cols + col - 1) % c.cols;
    int up = (r + 1) % c.rows, dn = (c.rows + r - 1) % c.rows;
    int* a = &c.adj[8 * k];
    a[0] = r * c.cols + cr;
    a[1] = r * c.cols + cl;
    a[2] = up * c.cols + col;
    a[3] = up * c.cols + cr;
    a[4] = up * c.cols + cl;
    a[5] = dn * c.cols + col;
    a[6] = dn * c.cols + cr;
    a[7] = dn * c.cols + cl;
  }
}

// find_cells, spherocyl_box.cpp:270-301
void bin_particles(orc_box* b) {
  Cells& c = b->cg;
  for (auto& m : c.members) m.clear();
  for (int p = 0; p < b->n; ++p) {
    wrap_pos(b->x[p], b->Lx);
    wrap_pos(b->y[p], b->Ly);
    wrap_angle(b->phi[p]);
    int col = (int)(b->x[p] / c.w);
    int row = (int)(b->y[p] / c.h);
    int id = col + row * c.cols;
    b->cell[p] = id;
    c.members[id].push_back(p);
  }
}

// check_neighbors, spherocyl_box.cpp:303-318
bool candidate(const orc_box* b, int p, int q) {
  double reach = b->R[p] + b->A[p] + b->R[q] + b->A[q];
  double dy = b->y[p] - b->y[q];
  dy += b->Ly * ((dy < -0.5 * b->Ly) - (dy > 0.5 * b->Ly));
  if (std::fabs(dy) < reach + b->pad) {
    double dx = b->x[p] - b->x[q];
    dx += b->Lx * ((dx < -0.5 * b->Lx) - (dx > 0.5 * b->Lx));
    // the largest strain before the remap: 0.5 (Lx / Ly), which is the
    // reference's 0.5 in its square boxes
    double dx_max_shear = dx + (0.5 * (b->Lx / b->Ly)) * dy;
    dx += b->gamma * dy;
    if (std::fabs(dx) < reach + b->pad || std::fabs(dx_max_shear) < reach + b->pad) return true;
  }
  return false;
}

// find_neighbors, spherocyl_box.cpp:320-355
void build_neighbors(orc_box* b) {
  bin_particles(b);
  for (int p = 0; p < b->n; ++p) {
    b->nbr[p].clear();
    b->xm[p] = 0.0;
    b->ym[p] = 0.0;
  }
  const Cells& c = b->cg;
  for (int p = 0; p < b->n - 1; ++p) {
    int home = b->cell[p];
    for (int q : c.members[home])
      if (q > p && candidate(b, p, q)) b->nbr[p].push_back(q);
    for (int k = 0; k < 8; ++k)
      for (int q : c.members[c.adj[8 * home + k]])
        if (q > p && candidate(b, p, q)) b->nbr[p].push_back(q);
  }
}

// Result of one segment-segment closest-approach evaluation.
struct PairOut {
  bool contact;
  double fx, fy, tau_p, tau_q;   // force on p (q gets -f) and torques
  double s, t, dx, dy;           // contact parameters and D vector
  double nxA, nyA, nxB, nyB;
  double dist, sigma, dV;
};

// Crossed segments (their lines meet inside both): the closest distance is 0
// and stays 0 under small displacements, and D is rounding noise (the
// reference normalises it -- a full-size force in a random direction -- or
// divides 0 by 0).  The overlap energy is continued to negative distance
// d = -delta, delta the minimal translation that uncrosses them: the smallest
// distance of an endpoint from the other segment's line.  That endpoint's
// segment is pushed back across the line with force (1 + delta/sigma)/sigma,
// applied at the endpoint and at its projection on the other segment; force
// and energy are continuous where the rods uncross.
void crossed_contact(double ddx, double ddy, PairOut& o) {
  const double nxA = o.nxA, nyA = o.nyA, nxB = o.nxB, nyB = o.nyB, sig = o.sigma;
  const double la2 = nxA * nxA + nyA * nyA, lb2 = nxB * nxB + nyB * nyB;
  const double la = std::sqrt(la2), lb = std::sqrt(lb2);
  const double mAx = -nyA / la, mAy = nxA / la, mBx = -nyB / lb, mBy = nxB / lb;
  const double hp = (ddx + nxA) * mBx + (ddy + nyA) * mBy;
  const double hm = (ddx - nxA) * mBx + (ddy - nyA) * mBy;
  const double gp = (-ddx + nxB) * mAx + (-ddy + nyB) * mAy;
  const double gm = (-ddx - nxB) * mAx + (-ddy - nyB) * mAy;
  const double dA = std::fmin(std::fabs(hp), std::fabs(hm));
  const double dB = std::fmin(std::fabs(gp), std::fabs(gm));
  double s, t, ux, uy, delta;
  if (dA <= dB) {
    s = std::fabs(hp) <= std::fabs(hm) ? 1.0 : -1.0;
    const double h = s > 0 ? hp : hm;
    ux = h > 0 ? -mBx : mBx;
    uy = h > 0 ? -mBy : mBy;
    t = std::fmin(std::fmax(((ddx + s * nxA) * nxB + (ddy + s * nyA) * nyB) / lb2, -1.0), 1.0);
    delta = dA;
  } else {
    t = std::fabs(gp) <= std::fabs(gm) ? 1.0 : -1.0;
    const double g = t > 0 ? gp : gm;
    ux = g > 0 ? mAx : -mAx;
    uy = g > 0 ? mAy : -mAy;
    s = std::fmin(std::fmax(((-ddx + t * nxB) * nxA + (-ddy + t * nyB) * nyA) / la2, -1.0), 1.0);
    delta = dB;
  }
  o.contact = true;
  o.dV = (1.0 + delta / sig) / sig;
  o.fx = ux * o.dV;
  o.fy = uy * o.dV;
  o.tau_p = s * nxA * o.fy - s * nyA * o.fx;
  o.tau_q = t * nyB * o.fx - t * nxB * o.fy;
  o.s = s;
  o.t = t;
  o.dist = -delta;
  o.dx = -delta * ux;
  o.dy = -delta * uy;
}

// Whether the segments intersect: each one's endpoints lie strictly on opposite
// sides of the other's line.  The closest-approach test of crossing (|tu| < 1,
// |s| < 1) also fires on rounding noise for nearly parallel segments (det =
// ac - b^2 cancels); a crossing is taken only when this confirms it.
bool segments_cross(double ddx, double ddy, double nxA, double nyA, double nxB, double nyB) {
  const double hp = nxB * (ddy + nyA) - nyB * (ddx + nxA);
  const double hm = nxB * (ddy - nyA) - nyB * (ddx - nxA);
  const double gp = nxA * (nyB - ddy) - nyA * (nxB - ddx);
  const double gm = nxA * (-nyB - ddy) - nyA * (-nxB - ddx);
  return ((hp > 0 && hm < 0) || (hp < 0 && hm > 0)) && ((gp > 0 && gm < 0) || (gp < 0 && gm > 0));
}

// calc_force_pair / calc_temp_force_pair / calc_se_force_pair geometry,
// strain_dynamics.cpp:17-75 (operation order kept for rounding agreement).
inline PairOut pair_eval(double xp, double yp, double php, double Rp, double Ap, double xq,
                         double yq, double phq, double Rq, double Aq, double Lx, double Ly,
                         double gam) {
  PairOut o;
  double ddx = xp - xq;
  double ddy = yp - yq;
  o.sigma = Rp + Rq;
  ddx += Lx * ((ddx < -0.5 * Lx) - (ddx > 0.5 * Lx));
  ddy += Ly * ((ddy < -0.5 * Ly) - (ddy > 0.5 * Ly));
  ddx += gam * ddy;
  o.nxA = Ap * std::cos(php);
  o.nyA = Ap * std::sin(php);
  o.nxB = Aq * std::cos(phq);
  o.nyB = Aq * std::sin(phq);
  double a = Ap * Ap;
  double bb = -(o.nxA * o.nxB + o.nyA * o.nyB);
  double c = Aq * Aq;
  double d = o.nxA * ddx + o.nyA * ddy;
  double e = -o.nxB * ddx - o.nyB * ddy;
  double det = a * c - bb * bb;
  double tu = (bb * d - a * e) / det;
  double t = std::fmin(std::fmax(tu, -1.), 1.);
  double s = -(bb * t + d) / a;
  double sabs = std::fabs(s);
  s = std::fmin(std::fmax(s, -1.), 1.);
  if (sabs > 1) t = std::fmin(std::fmax(-(bb * s + e) / c, -1.), 1.);
  o.s = s;
  o.t = t;
  o.dx = ddx + s * o.nxA - t * o.nxB;
  o.dy = ddy + s * o.nyA - t * o.nyB;
  double d2 = o.dx * o.dx + o.dy * o.dy;
  o.contact = d2 < o.sigma * o.sigma;
  if (o.contact && det > 0 && std::fabs(tu) < 1 && sabs < 1 &&
      segments_cross(ddx, ddy, o.nxA, o.nyA, o.nxB, o.nyB)) {  // crossed segments
    crossed_contact(ddx, ddy, o);
    return o;
  }
  if (o.contact && d2 == 0.0) {  // touching exactly (ends, collinear overlap): no direction
    o.dist = 0.0;
    o.dV = 1.0 / o.sigma;
    o.fx = o.fy = o.tau_p = o.tau_q = 0.0;
    return o;
  }
  if (o.contact) {
    o.dist = std::sqrt(d2);
    o.dV = (1.0 - o.dist / o.sigma) / o.sigma;
    o.fx = o.dx * o.dV / o.dist;
    o.fy = o.dy * o.dV / o.dist;
    o.tau_p = s * o.nxA * o.fy - s * o.nyA * o.fx;
    o.tau_q = t * o.nyB * o.fx - t * o.nxB * o.fy;
  } else {
    o.dist = o.dV = o.fx = o.fy = o.tau_p = o.tau_q = 0;
  }
  return o;
}

// Fills o's force fields from its D vector (harmonic contact, as pair_eval).
inline void contact_force(PairOut& o, double s, double t, bool crossed, double ddx, double ddy) {
  if (crossed) {  // as in pair_eval
    crossed_contact(ddx, ddy, o);
    return;
  }
  o.s = s;
  o.t = t;
  double d2 = o.dx * o.dx + o.dy * o.dy;
  o.contact = d2 < o.sigma * o.sigma;
  if (o.contact && d2 == 0.0) {
    o.dist = 0.0;
    o.dV = 1.0 / o.sigma;
    o.fx = o.fy = o.tau_p = o.tau_q = 0.0;
    return;
  }
  if (o.contact) {
    o.dist = std::sqrt(d2);
    o.dV = (1.0 - o.dist / o.sigma) / o.sigma;
    o.fx = o.dx * o.dV / o.dist;
    o.fy = o.dy * o.dV / o.dist;
    o.tau_p = s * o.nxA * o.fy - s * o.nyA * o.fx;
    o.tau_q = t * o.nyB * o.fx - t * o.nxB * o.fy;
  } else {
    o.dist = o.dV = o.fx = o.fy = o.tau_p = o.tau_q = 0;
  }
}

inline double clamp1(double v) { return std::fmin(std::fmax(v, -1.0), 1.0); }

// Two-point contact model of the reference's GPU code
// (src/cuda/2point_dynamics.cu:391-546): besides the closest pair of points
// (s1, t1), a second pair (s2, t2) -- for nearly parallel rods the other end of
// their overlap -- carries its own harmonic contact when it overlaps too, so the
// energy of two overlapping, nearly aligned rods has no kink at alignment.  The
// second pair starts from the opposite end of whichever segment set the first
// (s2 = -s1 or t2 = -t1), projected onto the other segment; a third candidate
// (s3, t3) from the other segment's matching end replaces it when closer.  The
// pair is evaluated once, lower id first, and both particles get +-F (the
// reference evaluates it from each side).  Returns the number of contacts in
// out[0..1] (out[0] is the closest-point contact of pair_eval).
inline int pair_eval_2p(double xp, double yp, double php, double Rp, double Ap, double xq,
                        double yq, double phq, double Rq, double Aq, double Lx, double Ly,
                        double gam, PairOut* out) {
  PairOut o;
  double ddx = xp - xq;
  double ddy = yp - yq;
  o.sigma = Rp + Rq;
  ddx += Lx * ((ddx < -0.5 * Lx) - (ddx > 0.5 * Lx));
  ddy += Ly * ((ddy < -0.5 * Ly) - (ddy > 0.5 * Ly));
  ddx += gam * ddy;
  o.nxA = Ap * std::cos(php);
  o.nyA = Ap * std::sin(php);
  o.nxB = Aq * std::cos(phq);
  o.nyB = Aq * std::sin(phq);
  const double a = Ap * Ap;
  const double b = -(o.nxA * o.nxB + o.nyA * o.nyB);
  const double c = Aq * Aq;
  const double d = o.nxA * ddx + o.nyA * ddy;
  const double e = -o.nxB * ddx - o.nyB * ddy;
  const double det = a * c - b * b;
  double s1, t1, s2, t2, s3, t3;
  bool crossed = false;
  if (det <= 0) {  // parallel: the two ends of the overlap
    s1 = clamp1(-(b + d) / a);
    s2 = clamp1(-(d - b) / a);
    if (std::fabs(s1) == 1) t1 = clamp1(-(b * s1 + e) / c);
    else t1 = b < 0 ? clamp1(-(b + e) / c) : clamp1(-(e - b) / c);
    if (std::fabs(s2) == 1) t2 = clamp1(-(b * s2 + e) / c);
    else t2 = b < 0 ? clamp1(-(e - b) / c) : clamp1(-(b + e) / c);
    s3 = s2;
    t3 = t2;
  } else {
    const double tu = (b * d - a * e) / det;
    t1 = clamp1(tu);
    s1 = -(b * t1 + d) / a;
    double sarg = std::fabs(s1);
    s1 = clamp1(s1);
    crossed = std::fabs(tu) < 1 && sarg < 1 && segments_cross(ddx, ddy, o.nxA, o.nyA, o.nxB, o.nyB);
    if (sarg >= 1) {
      t1 = clamp1(-(b * s1 + e) / c);
      s2 = -s1;
      t2 = -(b * s2 + e) / c;
      double targ = std::fabs(t2);
      t2 = clamp1(t2);
      if (targ > 1) {
        s2 = clamp1(-(b * t2 + d) / a);
        s3 = s2;
        t3 = t2;
      } else {
        t3 = b < 0 ? -s1 : s1;
        s3 = -(b * t3 + d) / a;
        sarg = std::fabs(s3);
        s3 = clamp1(s3);
        if (sarg >= 1) t3 = clamp1(-(b * s3 + e) / c);
      }
    } else {
      t2 = -t1;
      s2 = -(b * t2 + d) / a;
      sarg = std::fabs(s2);
      s2 = clamp1(s2);
      if (sarg > 1) {
        t2 = clamp1(-(b * s2 + e) / c);
        s3 = s2;
        t3 = t2;
      } else {
        s3 = b < 0 ? -t1 : t1;
        t3 = -(b * s3 + e) / c;
        double targ = std::fabs(t3);
        t3 = clamp1(t3);
        if (targ >= 1) s3 = clamp1(-(b * t3 + d) / a);
      }
    }
  }
  int k = 0;
  PairOut o1 = o;
  o1.dx = ddx + s1 * o.nxA - t1 * o.nxB;
  o1.dy = ddy + s1 * o.nyA - t1 * o.nyB;
  contact_force(o1, s1, t1, crossed, ddx, ddy);
  if (o1.contact) out[k++] = o1;
  if (s1 != s2 && t1 != t2) {
    PairOut o2 = o;
    o2.dx = ddx + s2 * o.nxA - t2 * o.nxB;
    o2.dy = ddy + s2 * o.nyA - t2 * o.nyB;
    double d22 = o2.dx * o2.dx + o2.dy * o2.dy;
    if (s3 != s1 && s3 != s2) {
      double dx3 = ddx + s3 * o.nxA - t3 * o.nxB;
      double dy3 = ddy + s3 * o.nyA - t3 * o.nyB;
      if (dx3 * dx3 + dy3 * dy3 < d22) {
        s2 = s3;
        t2 = t3;
        o2.dx = dx3;
        o2.dy = dy3;
      }
    }
    contact_force(o2, s2, t2, false, ddx, ddy);
    if (o2.contact) out[k++] = o2;
  }
  return k;
}

// Contacts of pair (p, q) under the box's contact model: out[0..k-1].
inline int pair_contacts(const orc_box* b, const double* x, const double* y, const double* phi,
                         int p, int q, double gam, double Lx, double Ly, PairOut* out) {
  if (b->contact_model == 1)
    return pair_eval_2p(x[p], y[p], phi[p], b->R[p], b->A[p], x[q], y[q], phi[q], b->R[q],
                        b->A[q], Lx, Ly, gam, out);
  out[0] = pair_eval(x[p], y[p], phi[p], b->R[p], b->A[p], x[q], y[q], phi[q], b->R[q], b->A[q],
                     Lx, Ly, gam);
  return out[0].contact ? 1 : 0;
}

inline PairOut pair_main(const orc_box* b, int p, int q, double gam, double Lx, double Ly) {
  return pair_eval(b->x[p], b->y[p], b->phi[p], b->R[p], b->A[p], b->x[q], b->y[q], b->phi[q],
                   b->R[q], b->A[q], Lx, Ly, gam);
}
inline PairOut pair_temp(const orc_box* b, int p, int q, double gam, double Lx, double Ly) {
  return pair_eval(b->tx[p], b->ty[p], b->tphi[p], b->R[p], b->A[p], b->tx[q], b->ty[q],
                   b->tphi[q], b->R[q], b->A[q], Lx, Ly, gam);
}

// calc_forces (strain_dynamics.cpp:137-153): zero, then half list with q > p.
void forces_main(orc_box* b) {
  for (int i = 0; i < b->n; ++i) b->fx[i] = b->fy[i] = b->tau[i] = 0.0;
  for (int p = 0; p < b->n; ++p)
    for (int q : b->nbr[p]) {
      if (q <= p) continue;
      PairOut pc[2];
      int k = pair_contacts(b, b->x.data(), b->y.data(), b->phi.data(), p, q, b->gamma, b->Lx,
                            b->Ly, pc);
      for (int j = 0; j < k; ++j) {
        const PairOut& o = pc[j];
        b->fx[p] += o.fx;
        b->fy[p] += o.fy;
        b->tau[p] += o.tau_p;
        b->fx[q] -= o.fx;
        b->fy[q] -= o.fy;
        b->tau[q] += o.tau_q;
      }
    }
}

// calc_temp_forces (strain_dynamics.cpp:156-172), with the predicted state and
// gamma + dt*gamma_dot; Lx/Ly are passed so the resize predictor can choose.
void forces_temp_at(orc_box* b, double Lx, double Ly, double gam);
void forces_temp(orc_box* b, double Lx, double Ly) {
  forces_temp_at(b, Lx, Ly, b->gamma + b->dt * b->strain_rate);
}
void forces_temp_at(orc_box* b, double Lx, double Ly, double gam) {
  for (int i = 0; i < b->n; ++i) b->tfx[i] = b->tfy[i] = b->ttau[i] = 0.0;
  for (int p = 0; p < b->n; ++p)
    for (int q : b->nbr[p]) {
      if (q <= p) continue;
      PairOut pc[2];
      int k = pair_contacts(b, b->tx.data(), b->ty.data(), b->tphi.data(), p, q, gam, Lx, Ly, pc);
      for (int j = 0; j < k; ++j) {
        const PairOut& o = pc[j];
        b->tfx[p] += o.fx;
        b->tfy[p] += o.fy;
        b->ttau[p] += o.tau_p;
        b->tfx[q] -= o.fx;
        b->tfy[q] -= o.fy;
        b->ttau[q] += o.tau_q;
      }
    }
}

// calc_se (spherocyl_box.cpp:425-522): forces plus energy and stress.  The
// reference's stress carries two slips (spherocyl_box.cpp:486, 493); E/Pxx/Pyy/Pxy
// reproduce them, cP** hold the corrected virial P_ab = sum F_b (CA_a - CB_a) / area.
void forces_se(orc_box* b) {
  for (int i = 0; i < b->n; ++i) b->fx[i] = b->fy[i] = b->tau[i] = 0.0;
  b->E = b->Pxx = b->Pyy = b->Pxy = 0.0;
  b->cPxx = b->cPyy = b->cPxy = b->cPyx = 0.0;
  b->ncontacts = 0;
  double area = b->Lx * b->Ly;
  for (int p = 0; p < b->n; ++p)
    for (int q : b->nbr[p]) {
      if (q <= p) continue;
      PairOut pc[2];
      int k = pair_contacts(b, b->x.data(), b->y.data(), b->phi.data(), p, q, b->gamma, b->Lx,
                            b->Ly, pc);
      if (k > 0) b->ncontacts++;
      for (int j = 0; j < k; ++j) {
        const PairOut& o = pc[j];
        b->fx[p] += o.fx;
        b->fy[p] += o.fy;
        b->tau[p] += o.tau_p;
        b->fx[q] -= o.fx;
        b->fy[q] -= o.fy;
        b->tau[q] += o.tau_q;
        double cxa = o.s * o.nxA - 0.5 * o.dx;
        double cya = o.s * o.nyA - 0.5 * o.dy;
        double cxb = o.t * o.nxB + 0.5 * o.dx;
        double cyb_ref = o.t * o.nxB + 0.5 * o.dy;  // reference slip (uses nxB)
        double cyb = o.t * o.nyB + 0.5 * o.dy;
        b->E += o.dV * o.sigma * (1.0 - o.dist / o.sigma) / (2.0 * b->n);
        b->Pxx += o.fx * cxa / area;
        b->Pxx -= o.fx * cxb / area;
        b->Pyy += o.fy * cya / area;
        b->Pyy -= o.fy * cyb_ref / area;
        b->Pxy += o.fy * cxa / area;
        b->Pxx -= o.fy * cxb / area;  // reference slip: should update Pxy
        b->cPxx += o.fx * (cxa - cxb) / area;
        b->cPyy += o.fy * (cya - cyb) / area;
        b->cPxy += o.fy * (cxa - cxb) / area;
        b->cPyx += o.fx * (cya - cyb) / area;
      }
    }
}

inline double iso_rot(double RC, double phi) { return 0.5 * (1 - RC * std::cos(2 * phi)); }

inline bool moved_too_far(const orc_box* b, int i) {
  return b->xm[i] > 0.5 * b->pad || b->xm[i] < -0.5 * b->pad || b->ym[i] > 0.5 * b->pad ||
         b->ym[i] < -0.5 * b->pad;
}

// set_back_gamma, spherocyl_box.cpp:524-545, in the general form of the GPU
// code for Lx != Ly (set_back_coords, src/cuda/find_neighbors.cu:320-330, and
// :369): x += (Lx/Ly) y and gamma -= Lx/Ly, a lattice translation of the
// sheared box.  With Lx = Ly (the CPU constructors) the ratio is exactly 1 and
// this is the CPU code's x += y, gamma -= 1; gamma_total is re-synchronised
// as the CPU code does only then.
void lees_edwards_remap(orc_box* b) {
  const double ratio = b->Lx / b->Ly;
  for (int i = 0; i < b->n; ++i) {
    if (b->y[i] < 0) b->y[i] += b->Ly;
    else if (b->y[i] > b->Ly) b->y[i] -= b->Ly;
    b->x[i] += ratio * b->y[i];
    if (b->x[i] < 0) b->x[i] += b->Lx;
    else
      while (b->x[i] > b->Lx) b->x[i] -= b->Lx;
  }
  b->gamma -= ratio;
  if (ratio == 1.0) b->gamma_total = int(b->gamma_total + 1) + b->gamma;
  build_neighbors(b);
}

// strain_step, strain_dynamics.cpp:179-235 (Heun; rebuild inside the loop as the
// reference does).
void heun_step(orc_box* b) {
  int n = b->n;
  std::vector<double> v0x(n), v0y(n), v0t(n);
  for (int i = 0; i < n; ++i) {
    v0x[i] = b->fx[i] - b->gamma * b->fy[i];
    v0y[i] = b->fy[i];
    v0t[i] = b->tau[i] / b->I[i] - b->strain_rate * iso_rot(b->RC[i], b->phi[i]);
    b->tx[i] = b->x[i] + b->dt * v0x[i];
    b->ty[i] = b->y[i] + b->dt * v0y[i];
    b->tphi[i] = b->phi[i] + b->dt * v0t[i];
  }
  forces_temp(b, b->Lx, b->Ly);
  double gam1 = b->gamma + b->dt * b->strain_rate;
  for (int i = 0; i < n; ++i) {
    double v1x = b->tfx[i] - gam1 * b->tfy[i];
    double v1y = b->tfy[i];
    double v1t = b->ttau[i] / b->I[i] - b->strain_rate * iso_rot(b->RC[i], b->tphi[i]);
    double ddx = 0.5 * b->dt * (v0x[i] + v1x);
    double ddy = 0.5 * b->dt * (v0y[i] + v1y);
    double ddp = 0.5 * b->dt * (v0t[i] + v1t);
    b->x[i] += ddx;
    b->y[i] += ddy;
    b->phi[i] += ddp;
    b->xm[i] += ddx;
    b->ym[i] += ddy;
    if (moved_too_far(b, i)) build_neighbors(b);
  }
  b->gamma += b->strain_rate * b->dt;
  b->gamma_total += b->strain_rate * b->dt;
  // the remap threshold 0.5 Lx/Ly (src/cuda/strain_dynamics.cu:712); 0.5 when Lx = Ly
  if (b->gamma >= 0.5 * (b->Lx / b->Ly)) lees_edwards_remap(b);
}

// Pure shear: pure_strain_step_2p of the reference's GPU code
// (src/cuda/2point_dynamics.cu:1663-2285; pure_euler_est_2p :1885-1890,
// pure_heun_corr_2p :2141-2166, the box :2273-2274) with the CPU code's
// conventions (mobility 1, torque / I).  x expands and y contracts at half the
// strain rate er:
//   predictor  x~ = x + dt (Fx + er x / 2),  y~ = y + dt (Fy - er y / 2),
//              phi~ = phi + dt (tau / I - er RC sin(2 phi) / 2)
//   box        Lx' = Lx + dt er Lx / 2,  Ly' = Ly - dt er Ly / 2
//   corrector  forces of the predicted state in the new box at the same
//              strain gamma; x = x~ + dt (Fx~ - Fx) / 2, likewise y, and phi
//              with torque / I
// gamma_total advances by dt er.  The oracle rebuilds cells and lists every
// step.
void pure_step(orc_box* b) {
  const int n = b->n;
  const double er = b->strain_rate, dt = b->dt;
  configure_cells(b);
  build_neighbors(b);
  forces_main(b);
  std::vector<double> f0x(n), f0y(n), f0t(n);
  for (int i = 0; i < n; ++i) {
    f0x[i] = b->fx[i];
    f0y[i] = b->fy[i];
    f0t[i] = b->tau[i] / b->I[i];
    b->tx[i] = b->x[i] + dt * (b->fx[i] + 0.5 * er * b->x[i]);
    b->ty[i] = b->y[i] + dt * (b->fy[i] - 0.5 * er * b->y[i]);
    b->tphi[i] = b->phi[i] + dt * (f0t[i] + er * (-0.5 * b->RC[i] * std::sin(2.0 * b->phi[i])));
  }
  const double Lxn = b->Lx + 0.5 * dt * er * b->Lx;
  const double Lyn = b->Ly - 0.5 * dt * er * b->Ly;
  forces_temp_at(b, Lxn, Lyn, b->gamma);
  for (int i = 0; i < n; ++i) {
    b->x[i] = b->tx[i] + 0.5 * dt * (b->tfx[i] - f0x[i]);
    b->y[i] = b->ty[i] + 0.5 * dt * (b->tfy[i] - f0y[i]);
    b->phi[i] = b->tphi[i] + 0.5 * dt * (b->ttau[i] / b->I[i] - f0t[i]);
  }
  b->Lx = Lxn;
  b->Ly = Lyn;
  b->gamma_total += dt * er;
  for (int i = 0; i < n; ++i) {
    wrap_pos(b->x[i], b->Lx);
    wrap_pos(b->y[i], b->Ly);
  }
  configure_cells(b);
  build_neighbors(b);
}

// flip_shear_direction (src/cuda/find_neighbors.cu:552-579)
void flip_shear(orc_box* b) {
  for (int i = 0; i < b->n; ++i) {
    b->x[i] = b->Lx - b->x[i];
    b->phi[i] = -b->phi[i];
  }
  b->gamma = -b->gamma;
  build_neighbors(b);
}

// reconfigure_cells, resize_dynamics.cpp:106-121 (the reference compares rows
// with Lx/hmin and cols with Ly/wmin; identical for Lx = Ly).
// rotate_by_gamma / gamma_rotate, src/cuda/find_neighbors.cu:580-645: the lab
// frame turns so that the sheared box vector (gamma Ly, Ly) is the new x axis.
void rotate_by_gamma(orc_box* b) {
  const double g = b->gamma, root = std::sqrt(1 + g * g);
  const double Lx = b->Lx, Ly = b->Ly;
  b->Lx = Ly * root;
  b->Ly = Lx / root;
  for (int i = 0; i < b->n; ++i) {
    const double x = b->x[i], y = b->y[i];
    double nx = (1 - 2 * std::signbit(g)) * y * root;
    double ny = (2 * std::signbit(g) - 1) * x / root;
    double np = g != 0.0 ? b->phi[i] - std::atan(1 / g) : b->phi[i] - kPi / 2;
    if (np > kPi) np -= 2 * kPi;
    else if (np < -kPi) np += 2 * kPi;
    if (ny > b->Ly) ny -= b->Ly;
    else if (ny < 0) ny += b->Ly;
    if (nx < 0) nx += b->Lx;
    else if (nx > b->Lx) nx -= b->Lx;
    b->x[i] = nx;
    b->y[i] = ny;
    b->phi[i] = np;
  }
  b->gamma = -g;
  configure_cells(b);
  build_neighbors(b);
}

void reconfigure_cells(orc_box* b) {
  double wmin = 2.24 * (b->Rmax + b->Amax) + b->pad;
  double hmin = 2 * (b->Rmax + b->Amax) + b->pad;
  int rows = std::max(static_cast<int>(b->Lx / hmin), 1);
  int cols = std::max(static_cast<int>(b->Ly / wmin), 1);
  if (rows != b->cg.rows || cols != b->cg.cols) {
    configure_cells(b);
    build_neighbors(b);
  } else {
    b->cg.w = b->Lx / b->cg.cols;
    b->cg.h = b->Ly / b->cg.rows;
  }
}

// resize_step, resize_dynamics.cpp:164-243.  The reference's predictor force
// pass uses the pre-step box length for the minimum image (corrected_image=0);
// corrected_image=1 uses the rescaled length instead.
void resize_step(orc_box* b, bool shrink, bool corrected_image) {
  int n = b->n;
  std::vector<double> f0x(n), f0y(n), t0(n);
  double scale = std::pow(1.0 - b->resize_rate, b->dt);
  for (int i = 0; i < n; ++i) {
    f0x[i] = b->fx[i];
    f0y[i] = b->fy[i];
    t0[i] = b->tau[i];
    b->tx[i] = b->x[i] + b->dt * f0x[i];
    b->ty[i] = b->y[i] + b->dt * f0y[i];
    b->tphi[i] = b->phi[i] + b->dt * t0[i] / b->I[i];
    if (shrink) {
      b->tx[i] *= std::pow(1.0 - b->resize_rate, b->dt);
      b->ty[i] *= std::pow(1.0 - b->resize_rate, b->dt);
    } else {
      b->tx[i] /= std::pow(1.0 - b->resize_rate, b->dt);
      b->ty[i] /= std::pow(1.0 - b->resize_rate, b->dt);
    }
  }
  if (corrected_image) {
    double Lx = shrink ? b->Lx * scale : b->Lx / scale;
    double Ly = shrink ? b->Ly * scale : b->Ly / scale;
    forces_temp(b, Lx, Ly);
  } else {
    forces_temp(b, b->Lx, b->Ly);
  }
  for (int i = 0; i < n; ++i) {
    double ddx = 0.5 * b->dt * (f0x[i] + b->tfx[i]);
    double ddy = 0.5 * b->dt * (f0y[i] + b->tfy[i]);
    double ddp = 0.5 * b->dt * (t0[i] + b->ttau[i]) / b->I[i];
    double x0 = b->x[i], y0 = b->y[i];
    b->x[i] += ddx;
    b->y[i] += ddy;
    b->phi[i] += ddp;
    if (shrink) {
      b->x[i] *= std::pow(1.0 - b->resize_rate, b->dt);
      b->y[i] *= std::pow(1.0 - b->resize_rate, b->dt);
    } else {
      b->x[i] /= std::pow(1.0 - b->resize_rate, b->dt);
      b->y[i] /= std::pow(1.0 - b->resize_rate, b->dt);
    }
    b->xm[i] += b->x[i] - x0;
    b->ym[i] += b->y[i] - y0;
    reconfigure_cells(b);
    if (moved_too_far(b, i)) build_neighbors(b);
  }
  if (shrink) {
    b->Lx *= std::pow(1.0 - b->resize_rate, b->dt);
    b->Ly *= std::pow(1.0 - b->resize_rate, b->dt);
  } else {
    b->Lx /= std::pow(1.0 - b->resize_rate, b->dt);
    b->Ly /= std::pow(1.0 - b->resize_rate, b->dt);
  }
}

// gradient_relax_step, relaxation_dynamics.cpp:11-31
void gd_step(orc_box* b) {
  for (int i = 0; i < b->n; ++i) {
    double ddx = b->dt * b->fx[i];
    double ddy = b->dt * b->fy[i];
    b->x[i] += ddx;
    b->y[i] += ddy;
    b->phi[i] += b->dt * b->tau[i] / b->I[i];
    b->xm[i] += ddx;
    b->ym[i] += ddy;
    if (moved_too_far(b, i)) build_neighbors(b);
  }
}

double residual_rms(const orc_box* b) {
  double s = 0;
  for (int i = 0; i < b->n; ++i) s += b->fx[i] * b->fx[i] + b->fy[i] * b->fy[i] + b->tau[i] * b->tau[i];
  return std::sqrt(s / (3.0 * b->n));
}

// ---------------------------------------------------------------------------
// FIRE -- the project's specification (no reference counterpart; Bitzek et al.,
// PRL 97, 170201 (2006), with the half-step retreat of FIRE 2.0, Guenole et al.,
// Comput. Mater. Sci. 175, 109584 (2020)).  Generalised coordinates per particle
// are (x, y, psi = phi*sqrt(I)) so that all three have unit mass; the generalised
// force is g = (Fx, Fy, tau/sqrt(I)).  One iteration from (x_k, v_k):
//   1. forces at x_k; stop if sqrt(sum(Fx^2+Fy^2+tau^2)/(3N)) < ftol
//   2. P = g.v, G = g.g, V = v.v (sums over all particles)
//   3. P >= 0: v <- (1-alpha) v + alpha sqrt(V/G) g; npos += 1;
//             if npos > n_min: dt <- min(f_inc dt, dt_max), alpha <- f_alpha alpha
//      else : x <- x - dt v / 2; v <- 0; dt <- f_dec dt; alpha <- alpha0; npos <- 0
//   4. v <- v + dt g; displacement = dt v, its (x, y) part scaled down to length
//      max_move if longer; x <- x + displacement
// The sums are taken in FP64.
struct FireParams {
  double dt0, dt_max, n_min, f_inc, f_dec, alpha0, f_alpha, max_move, ftol, max_iter;
};

FireParams fire_params(const double* p) {
  return FireParams{p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9]};
}

// Returns 1 when converged at entry (state untouched), 0 otherwise.
int fire_iteration(orc_box* b, const FireParams& fp, double* vx, double* vy, double* vp,
                   double* st) {
  forces_main(b);
  if (residual_rms(b) < fp.ftol) {
    st[3] = 1;
    return 1;
  }
  double P = 0, G = 0, V = 0;
  for (int i = 0; i < b->n; ++i) {
    double sq = std::sqrt(b->I[i]);
    double gx = b->fx[i], gy = b->fy[i], gp = b->tau[i] / sq;
    P += gx * vx[i] + gy * vy[i] + gp * vp[i];
    G += gx * gx + gy * gy + gp * gp;
    V += vx[i] * vx[i] + vy[i] * vy[i] + vp[i] * vp[i];
  }
  double dt = st[0], alpha = st[1], npos = st[2];
  bool mix = P >= 0;  // v = 0 at the start gives P = 0: keep going, do not shrink dt
  double c_keep = 0, c_force = 0, dt_new;
  if (mix) {
    c_keep = 1.0 - alpha;
    c_force = (G > 0) ? alpha * std::sqrt(V / G) : 0.0;
    npos += 1;
    if (npos > fp.n_min) {
      dt_new = std::fmin(dt * fp.f_inc, fp.dt_max);
      alpha = alpha * fp.f_alpha;
    } else {
      dt_new = dt;
    }
  } else {
    dt_new = dt * fp.f_dec;
    alpha = fp.alpha0;
    npos = 0;
  }
  for (int i = 0; i < b->n; ++i) {
    double sq = std::sqrt(b->I[i]);
    double gx = b->fx[i], gy = b->fy[i], gp = b->tau[i] / sq;
    double ux, uy, up;
    if (mix) {
      ux = c_keep * vx[i] + c_force * gx;
      uy = c_keep * vy[i] + c_force * gy;
      up = c_keep * vp[i] + c_force * gp;
    } else {
      b->x[i] -= 0.5 * dt * vx[i];
      b->y[i] -= 0.5 * dt * vy[i];
      b->phi[i] -= 0.5 * dt * vp[i] / sq;
      b->xm[i] -= 0.5 * dt * vx[i];
      b->ym[i] -= 0.5 * dt * vy[i];
      ux = uy = up = 0.0;
    }
    ux += dt_new * gx;
    uy += dt_new * gy;
    up += dt_new * gp;
    vx[i] = ux;
    vy[i] = uy;
    vp[i] = up;
    double mx = dt_new * ux, my = dt_new * uy;
    double m2 = mx * mx + my * my;
    if (m2 > fp.max_move * fp.max_move) {
      double sc = fp.max_move / std::sqrt(m2);
      mx *= sc;
      my *= sc;
    }
    b->x[i] += mx;
    b->y[i] += my;
    b->phi[i] += dt_new * up / sq;
    b->xm[i] += mx;
    b->ym[i] += my;
  }
  for (int i = 0; i < b->n; ++i)
    if (moved_too_far(b, i)) {
      build_neighbors(b);
      break;
    }
  st[0] = dt_new;
  st[1] = alpha;
  st[2] = npos;
  st[3] = 0;
  return 0;
}

}  // namespace

extern "C" {

orc_box* orc_new(double L, int n, const double* x, const double* y, const double* phi,
                 const double* A, const double* R, double pad) {
  orc_box* b = new orc_box();
  b->n = n;
  b->Lx = b->Ly = L;
  auto sz = [n](std::vector<double>& v) { v.assign(n, 0.0); };
  for (auto* v : {&b->x, &b->y, &b->phi, &b->R, &b->A, &b->I, &b->RC, &b->fx, &b->fy, &b->tau,
                  &b->xm, &b->ym, &b->tx, &b->ty, &b->tphi, &b->tfx, &b->tfy, &b->ttau})
    sz(*v);
  b->cell.assign(n, 0);
  b->nbr.assign(n, {});
  for (int p = 0; p < n; ++p) {
    b->x[p] = x[p];
    b->y[p] = y[p];
    b->phi[p] = phi[p];
    b->A[p] = A[p];
    b->R[p] = R[p];
    rot_consts(R[p], A[p], &b->I[p], &b->RC[p]);
    wrap_pos(b->x[p], b->Lx);
    wrap_pos(b->y[p], b->Ly);
    wrap_angle(b->phi[p]);
    if (A[p] > b->Amax) b->Amax = A[p];
    if (R[p] > b->Rmax) b->Rmax = R[p];
  }
  double filled = 0.0;  // calculate_packing_fraction, spherocyl_box.cpp:357-366
  for (int p = 0; p < n; ++p) filled += (kPi * b->R[p] + 4.0 * b->A[p]) * b->R[p];
  b->packing = filled / (b->Lx * b->Ly);
  b->pad = pad;
  configure_cells(b);
  build_neighbors(b);
  return b;
}

void orc_free(orc_box* b) { delete b; }

void orc_set_contact_model(orc_box* b, int model) { b->contact_model = model == 1 ? 1 : 0; }

void orc_set_params(orc_box* b, double gamma, double gamma_total, double strain_rate,
                    double resize_rate, double step) {
  b->gamma = gamma;
  b->gamma_total = gamma_total;
  b->strain_rate = strain_rate;
  b->resize_rate = resize_rate;
  b->dt = step;
}

void orc_set_positions(orc_box* b, const double* x, const double* y, const double* phi) {
  for (int i = 0; i < b->n; ++i) {
    if (x) b->x[i] = x[i];
    if (y) b->y[i] = y[i];
    if (phi) b->phi[i] = phi[i];
  }
}

void orc_set_box(orc_box* b, double Lx, double Ly) {
  b->Lx = Lx;
  b->Ly = Ly;
  configure_cells(b);
  build_neighbors(b);
}

void orc_find_neighbors(orc_box* b) { build_neighbors(b); }
void orc_calc_forces(orc_box* b) { forces_main(b); }
void orc_calc_se(orc_box* b) { forces_se(b); }
void orc_strain_step(orc_box* b) { heun_step(b); }
void orc_resize_step(orc_box* b, int shrink, int corrected_image) {
  resize_step(b, shrink != 0, corrected_image != 0);
}
void orc_gd_step(orc_box* b) { gd_step(b); }
void orc_pure_step(orc_box* b) { pure_step(b); }
void orc_flip_shear(orc_box* b) { flip_shear(b); }
void orc_rotate_by_gamma(orc_box* b) { rotate_by_gamma(b); }
void orc_set_back_gamma(orc_box* b) { lees_edwards_remap(b); }

int orc_run_steps(orc_box* b, int nsteps, int se_every, double* se_out) {
  int rec = 0;
  for (int s = 1; s <= nsteps; ++s) {
    heun_step(b);
    if (se_every > 0 && s % se_every == 0) {
      forces_se(b);
      if (se_out) {
        se_out[4 * rec + 0] = b->E;
        se_out[4 * rec + 1] = b->Pxx;
        se_out[4 * rec + 2] = b->Pyy;
        se_out[4 * rec + 3] = b->Pxy;
      }
      ++rec;
    } else {
      forces_main(b);
    }
  }
  return rec;
}

void orc_get_particles(const orc_box* b, double* x, double* y, double* phi, double* R, double* A,
                       double* fx, double* fy, double* tau, double* I, double* RC) {
  size_t bytes = sizeof(double) * b->n;
  if (x) std::memcpy(x, b->x.data(), bytes);
  if (y) std::memcpy(y, b->y.data(), bytes);
  if (phi) std::memcpy(phi, b->phi.data(), bytes);
  if (R) std::memcpy(R, b->R.data(), bytes);
  if (A) std::memcpy(A, b->A.data(), bytes);
  if (fx) std::memcpy(fx, b->fx.data(), bytes);
  if (fy) std::memcpy(fy, b->fy.data(), bytes);
  if (tau) std::memcpy(tau, b->tau.data(), bytes);
  if (I) std::memcpy(I, b->I.data(), bytes);
  if (RC) std::memcpy(RC, b->RC.data(), bytes);
}

void orc_force_scale(orc_box* b, double* sx, double* sy, double* st) {
  for (int i = 0; i < b->n; ++i) sx[i] = sy[i] = st[i] = 0.0;
  for (int p = 0; p < b->n; ++p)
    for (int q : b->nbr[p]) {
      if (q <= p) continue;
      PairOut o = pair_main(b, p, q, b->gamma, b->Lx, b->Ly);
      if (!o.contact) continue;
      sx[p] += std::fabs(o.fx);
      sy[p] += std::fabs(o.fy);
      st[p] += std::fabs(o.tau_p);
      sx[q] += std::fabs(o.fx);
      sy[q] += std::fabs(o.fy);
      st[q] += std::fabs(o.tau_q);
    }
}

void orc_get_scalars(const orc_box* b, double* out) {
  double v[20] = {b->Lx,  b->Ly,  b->gamma, b->gamma_total, b->strain_rate, b->resize_rate,
                  b->dt,  b->packing, b->pad, b->E, b->Pxx, b->Pyy, b->Pxy, b->cPxx, b->cPyy,
                  b->cPxy, b->cPyx, (double)b->cg.rows, (double)b->cg.cols, (double)b->ncontacts};
  std::memcpy(out, v, sizeof(v));
}

int orc_get_neighbors(const orc_box* b, int* p, int* q, int cap) {
  int k = 0;
  for (int i = 0; i < b->n; ++i)
    for (int j : b->nbr[i]) {
      if (k < cap) {
        p[k] = i;
        q[k] = j;
      }
      ++k;
    }
  return k;
}

int orc_get_contacts(orc_box* b, int* p, int* q, int cap) {
  int k = 0;
  for (int i = 0; i < b->n; ++i)
    for (int j : b->nbr[i]) {
      if (j <= i) continue;
      if (!pair_main(b, i, j, b->gamma, b->Lx, b->Ly).contact) continue;
      if (k < cap) {
        p[k] = i;
        q[k] = j;
      }
      ++k;
    }
  return k;
}

int orc_brute_contacts(orc_box* b, int* p, int* q, int cap) {
  int k = 0;
  for (int i = 0; i < b->n; ++i)
    for (int j = i + 1; j < b->n; ++j) {
      if (!pair_main(b, i, j, b->gamma, b->Lx, b->Ly).contact) continue;
      if (k < cap) {
        p[k] = i;
        q[k] = j;
      }
      ++k;
    }
  return k;
}

// Contacts of one pair (p = A, q = B) under contact model `model` in an open
// square box of side L at strain gamma: out[9 k .. 9 k + 8] = s t Dx Dy dist fx
// fy tau_p tau_q of contact k (the force acts on p).  Returns the count.
int orc_pair_contacts(int model, double xp, double yp, double php, double Rp, double Ap,
                      double xq, double yq, double phq, double Rq, double Aq, double L,
                      double gamma, double* out) {
  PairOut pc[2];
  int k;
  if (model == 1) {
    k = pair_eval_2p(xp, yp, php, Rp, Ap, xq, yq, phq, Rq, Aq, L, L, gamma, pc);
  } else {
    pc[0] = pair_eval(xp, yp, php, Rp, Ap, xq, yq, phq, Rq, Aq, L, L, gamma);
    k = pc[0].contact ? 1 : 0;
  }
  for (int j = 0; j < k; ++j) {
    const PairOut& o = pc[j];
    const double v[9] = {o.s, o.t, o.dx, o.dy, o.dist, o.fx, o.fy, o.tau_p, o.tau_q};
    for (int i = 0; i < 9; ++i) out[9 * j + i] = v[i];
  }
  return k;
}

double orc_frms(orc_box* b) {
  forces_main(b);
  return residual_rms(b);
}

void orc_fire_iter(orc_box* b, const double* params, double* vx, double* vy, double* vpsi,
                   double* fire_state) {
  FireParams fp = fire_params(params);
  fire_iteration(b, fp, vx, vy, vpsi, fire_state);
}

int orc_fire(orc_box* b, const double* params, double* frms_out) {
  FireParams fp = fire_params(params);
  std::vector<double> vx(b->n, 0.0), vy(b->n, 0.0), vp(b->n, 0.0);
  double st[4] = {fp.dt0, fp.alpha0, 0.0, 0.0};
  int it = 0;
  for (; it < (int)fp.max_iter; ++it)
    if (fire_iteration(b, fp, vx.data(), vy.data(), vp.data(), st)) break;
  forces_main(b);
  if (frms_out) *frms_out = residual_rms(b);
  return it;
}

}  // extern "C"
