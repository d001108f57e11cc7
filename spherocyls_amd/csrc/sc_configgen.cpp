// Seeded random configuration generator (host code, C ABI).
//
// Counterpart of SpherocylBox::random_configuration (src/cpp/spherocyl_box.cpp:694-790):
// particles are placed one after another at uniform random centres and angles,
// and a placement is redrawn while it is rejected against an already placed
// particle.  RANDOM / RANDOM_ALIGNED reject a "crossing", i.e. a squared closest
// segment distance below 1e-5 (R_p + R_q)^2 or NaN (check_particle_cross,
// :597-645): overlaps are allowed, exact crossings are not, since a crossing
// gives a zero-length contact vector whose force direction is undefined (0/0 in
// the pair force).  ZERO_E modes and the grid modes reject any contact
// (check_particle_contact, :547-595).  ALIGNED modes set every angle to 0.
// Unlike the reference (srand(time(0)), :739) the stream is seeded.
#include <cmath>
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/spherocyls.h"
#include "sc_status.h"

namespace {

// SplitMix64-seeded xoshiro256** generator (Blackman & Vigna, 2018).
struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& z) {
    z += 0x9e3779b97f4a7c15ull;
    uint64_t r = z;
    r = (r ^ (r >> 30)) * 0xbf58476d1ce4e5b9ull;
    r = (r ^ (r >> 27)) * 0x94d049bb133111ebull;
    return r ^ (r >> 31);
  }
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& v : s) v = splitmix(z);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

// Standard normal variate (Box-Muller transform of two uniforms).
double normal(Rng& rng) {
  const double kTwoPi = 6.283185307179586476925287;
  double u1 = rng.uniform();
  while (u1 <= 0.0) u1 = rng.uniform();
  const double u2 = rng.uniform();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(kTwoPi * u2);
}

// Gamma(shape k, scale theta) variate by Marsaglia and Tsang's squeeze method
// (G. Marsaglia, W. W. Tsang, "A simple method for generating gamma
// variables", ACM Trans. Math. Softw. 26 (2000) 363-372); for k < 1 the
// variate of shape k + 1 times u^(1/k).
double gamma_variate(Rng& rng, double k, double theta) {
  if (k < 1.0) {
    double u = rng.uniform();
    while (u <= 0.0) u = rng.uniform();
    return gamma_variate(rng, k + 1.0, theta) * std::pow(u, 1.0 / k);
  }
  const double d = k - 1.0 / 3.0, c = 1.0 / std::sqrt(9.0 * d);
  for (;;) {
    const double x = normal(rng);
    double v = 1.0 + c * x;
    if (v <= 0.0) continue;
    v = v * v * v;
    const double u = rng.uniform();
    if (u < 1.0 - 0.0331 * x * x * x * x) return d * v * theta;
    if (u > 0.0 && std::log(u) < 0.5 * x * x + d * (1.0 - v + std::log(v))) return d * v * theta;
  }
}

// Squared closest distance between two segments (same closest-approach
// parametrisation as the pair force) for positions in sheared coordinates at
// strain gamma: minimum image in each coordinate, then dx += gamma dy, the pair
// separation of src/cpp/strain_dynamics.cpp:27-31.
double seg_dist2(double xp, double yp, double pp, double Ap, double xq, double yq, double pq,
                 double Aq, double L, double gamma) {
  double dx = xp - xq, dy = yp - yq;
  dx += L * ((dx < -0.5 * L) - (dx > 0.5 * L));
  dy += L * ((dy < -0.5 * L) - (dy > 0.5 * L));
  dx += gamma * dy;
  double nxa = Ap * std::cos(pp), nya = Ap * std::sin(pp);
  double nxb = Aq * std::cos(pq), nyb = Aq * std::sin(pq);
  double a = Ap * Ap, b = -(nxa * nxb + nya * nyb), c = Aq * Aq;
  double d = nxa * dx + nya * dy, e = -nxb * dx - nyb * dy;
  double det = a * c - b * b;
  double t = std::fmin(std::fmax((b * d - a * e) / det, -1.), 1.);
  double s = -(b * t + d) / a;
  double sa = std::fabs(s);
  s = std::fmin(std::fmax(s, -1.), 1.);
  if (sa > 1) t = std::fmin(std::fmax(-(b * s + e) / c, -1.), 1.);
  double ex = dx + s * nxa - t * nxb, ey = dy + s * nya - t * nyb;
  return ex * ex + ey * ey;
}

}  // namespace

namespace {

int random_configuration(int n, double L, double gamma, const double* A, const double* R,
                         int config, unsigned long long seed, int max_tries, double* x, double* y,
                         double* phi) {
  if (n <= 0 || !(L > 0) || !(std::fabs(gamma) <= 0.5) || !A || !R || !x || !y || !phi ||
      config < SC_CONFIG_RANDOM || config > SC_CONFIG_GRID_ALIGNED || max_tries < 1)
    return SC_ERR_ARG;
  double reach = 0, amax = 0, rmax = 0;
  for (int i = 0; i < n; ++i) {
    if (!(R[i] > 0) || !(A[i] >= 0)) return SC_ERR_ARG;
    reach = std::fmax(reach, 2 * (R[i] + A[i]));
    amax = std::fmax(amax, A[i]);
    rmax = std::fmax(rmax, R[i]);
  }
  const bool aligned = config % 2 == 1;
  const bool no_contact = config >= SC_CONFIG_ZERO_E;  // zero-energy and grid modes
  const bool on_grid = config >= SC_CONFIG_GRID;
  // Grid modes (reference config > 3): one particle per rectangular site, rows
  // and columns in the ratio of the particle's aspect (A+R)/R, each centre
  // drawn inside its site away from the site edges by (A+R, R).
  int grows = 1, gcols = 1;
  double gw = 0, gh = 0, gxo = amax + rmax, gyo = rmax;
  if (on_grid) {
    double aspect = (amax + rmax) / rmax;
    double dcols = std::sqrt(n / aspect);
    grows = (int)(aspect * dcols);
    gcols = (int)dcols;
    while ((long long)grows * gcols < n) ++grows;  // every particle gets a site
    gw = L / gcols;
    gh = L / grows;
    if (gw < 2 * (amax + rmax) || gh < 2 * rmax) return SC_ERR_PLACEMENT;
  }
  // Uniform search grid in sheared coordinates; with |gamma| <= 0.5 a pair within
  // one reach (lab frame) is at most 1.12 reach apart in sheared x.
  int nc = std::max(1, (int)(L / (1.12 * reach)));
  double w = L / nc;
  std::vector<std::vector<int>> grid((size_t)nc * nc);
  Rng rng(seed);
  const double kPi = 3.14159265358979323846;
  auto rejects = [&](int p, double px, double py, double pp, int q) {
    double sig = R[p] + R[q];
    double d2 = seg_dist2(px, py, pp, A[p], x[q], y[q], phi[q], A[q], L, gamma);
    double lim = no_contact ? sig * sig : 1e-5 * sig * sig;
    return d2 < lim || d2 != d2;
  };
  for (int p = 0; p < n; ++p) {
    int tries = 0;
    for (;;) {
      if (++tries > max_tries) return SC_ERR_PLACEMENT;
      double px, py;
      if (on_grid) {
        int gr = p / gcols, gc = p % gcols;
        px = gc * gw + gxo + (gw - 2 * gxo) * rng.uniform();
        py = gr * gh + gyo + (gh - 2 * gyo) * rng.uniform();
      } else {
        px = L * rng.uniform();
        py = L * rng.uniform();
      }
      double pp = aligned ? 0.0 : kPi * rng.uniform();
      int cx = std::min(nc - 1, (int)(px / w)), cy = std::min(nc - 1, (int)(py / w));
      bool bad = false;
      if (nc < 3) {  // small box: compare with everybody
        for (int q = 0; q < p && !bad; ++q) bad = rejects(p, px, py, pp, q);
      } else {
        for (int oy = -1; oy <= 1 && !bad; ++oy)
          for (int ox = -1; ox <= 1 && !bad; ++ox) {
            int gx = (cx + ox + nc) % nc, gy = (cy + oy + nc) % nc;
            for (int q : grid[(size_t)gy * nc + gx])
              if (rejects(p, px, py, pp, q)) {
                bad = true;
                break;
              }
          }
      }
      if (bad) continue;
      x[p] = px;
      y[p] = py;
      phi[p] = pp;
      grid[(size_t)cy * nc + cx].push_back(p);
      break;
    }
  }
  return SC_OK;
}

}  // namespace

extern "C" int sc_random_configuration(int n, double L, double gamma, const double* A,
                                       const double* R, int config, unsigned long long seed,
                                       int max_tries, double* x, double* y, double* phi) {
  try {
    return random_configuration(n, L, gamma, A, R, config, seed, max_tries, x, y, phi);
  } catch (...) {
    return sc::current_exception_status(nullptr);
  }
}

extern "C" int sc_gamma_shapes(int n, double aspect_mean, double aspect_rel_sigma,
                               double bidispersity, unsigned long long seed, double* R,
                               double* A) {
  if (n <= 0 || !(aspect_mean > 0) || !(aspect_rel_sigma > 0) || !(bidispersity >= 1) ||
      !std::isfinite(aspect_mean) || !std::isfinite(aspect_rel_sigma) || !R || !A)
    return SC_ERR_ARG;
  try {
    // set_shapes with m_dASig > 0 (src/cuda/spherocyl_box.cpp:1266-1300):
    // A/R ~ Gamma(k = mu^2 / sigma^2, theta = sigma^2 / mu) with
    // sigma = aspect_rel_sigma * mu, and R = 0.5 bidispersity for even
    // indices, 0.5 for odd ones
    const double sig = aspect_mean * aspect_rel_sigma;
    const double k = aspect_mean * aspect_mean / (sig * sig), theta = sig * sig / aspect_mean;
    Rng rng(seed);
    for (int p = 0; p < n; ++p) {
      R[p] = 0.5 + (1 - (p % 2)) * (bidispersity - 1.0) * 0.5;
      A[p] = R[p] * gamma_variate(rng, k, theta);
    }
    return SC_OK;
  } catch (...) {
    return sc::current_exception_status(nullptr);
  }
}

extern "C" int sc_random_packing(int n, double L, double gamma, const double* A, const double* R,
                                 unsigned long long seed, int max_tries, double* x, double* y,
                                 double* phi) {
  return sc_random_configuration(n, L, gamma, A, R, SC_CONFIG_RANDOM, seed, max_tries, x, y, phi);
}
