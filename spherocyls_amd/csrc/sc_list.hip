// Neighbour-list force kernels (one thread per particle).
//
// Latency, not bandwidth or FP64 rate, bounds the tile kernel: every tile walks
// three dependent global-load rounds and three barriers before it computes.  Here
// a particle's partners are kept in a column-major list ([k][ntot], coalesced
// per k), chosen with a skin, and gathered straight from the cell-sorted
// particle arrays, so a wavefront never waits at a barrier and thousands of
// independent particles hide the gather latency.
//
// Lists are rebuilt lazily inside the force pass: the kernel that produces a
// state (predictor, corrector, FIRE/GD update) raises its packing's flag when a
// particle has moved more than skin/2 (per sheared coordinate) from its position
// at the last build, when a Lees-Edwards remap happened, or when the strain
// leaves the window the lists were built for.  The next force pass then rebuilds
// every row of that packing from the exact per-pass cell buckets, each thread its
// own row, before using it; no host decision and no extra launch are involved.
//
// Selection rule (skin s <= cell pad; S = R_p + A_p + R_q + A_q; dx_s, dy minimum
// images in sheared coordinates; strain window [g_lo, g_hi]; c = max(1.5,
// 1 + Lx / (2 Ly))):
//     |dy| < S + s  and  ( |dx_s + g_lo dy| < S + c s  or  |dx_s + g_hi dy| < S + c s )
// A pair that touches at any strain in the window with every particle within
// s/2 of its build position satisfies it: its lab x-separation at that strain
// changed by less than (1 + |g|) s <= c s (the strain stays below the remap
// threshold Lx / (2 Ly), which is 1/2 in a square box), and a linear function
// of the strain is smallest in magnitude at an end of the window unless it
// crosses zero, when one end is within |dy| (g_hi - g_lo) / 2.  Such a pair
// also lies in adjacent cells at build time (the cell width bounds its
// sheared x-separation).
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_physics.h"

namespace sc {

// Ascending bitonic sort by key of the first W (key, value) pairs held in
// registers (every index is a compile-time constant once the loops unroll).
template <int W>
__device__ __forceinline__ void sort_pairs(int (&key)[kCapMin], int (&val)[kCapMin]) {
  static_assert((W & (W - 1)) == 0 && W <= kCapMin, "width must be a power of two <= kCapMin");
#pragma unroll
  for (int k = 2; k <= W; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < W; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const bool swap = up ? key[i] > key[l] : key[i] < key[l];
          const int ki = key[i], vi = val[i];
          key[i] = swap ? key[l] : ki;
          val[i] = swap ? val[l] : vi;
          key[l] = swap ? ki : key[l];
          val[l] = swap ? vi : val[l];
        }
      }
}

// Append the members of one bucket (m <= W of them) that pass the list's
// selection rule to particle g's row, in increasing particle id.  The bucket
// is read once (its slots share one 64-byte line) and sorted in registers.
template <int W>
__device__ __forceinline__ void row_from_bucket(const ListArgs& a, const int* sl, int m, int g,
                                                double xi, double yi, double reach_i, double Lx,
                                                double Ly, double s, double sx, double glo,
                                                double ghi, int& n) {
#pragma clang fp contract(off)
  int val[kCapMin], key[kCapMin];
#pragma unroll
  for (int k = 0; k < W; ++k) val[k] = k < m ? sl[k] : g;
#pragma unroll
  for (int k = 0; k < W; ++k) key[k] = k < m ? a.b.id[val[k]] : 0x7fffffff;
  sort_pairs<W>(key, val);
#pragma unroll
  for (int k = 0; k < W; ++k) {
    const int q = val[k];
    if (k < m && q != g) {
      const double reach = reach_i + a.b.R[q] + a.b.A[q];
      const double dy = min_image(yi - a.ey[q], Ly);
      const double dx = min_image(xi - a.ex[q], Lx);
      const double flo = dx + glo * dy, fhi = dx + ghi * dy;
      if (fabs(dy) < reach + s && (fabs(flo) < reach + sx || fabs(fhi) < reach + sx)) {
        if (n < kNbrCap) a.nbr[(size_t)n * a.ntot + g] = q;
        ++n;
      }
    }
  }
}

// A bucket of more than kCapMin members (crowded cells of polydisperse
// packings): the members in increasing id by repeated search for the next
// larger id, each tested as row_from_bucket does.
__device__ __forceinline__ void row_from_crowded_bucket(const ListArgs& a, const int* sl, int m,
                                                     int g, double xi, double yi, double reach_i,
                                                     double Lx, double Ly, double s, double sx,
                                                     double glo, double ghi, int& n) {
#pragma clang fp contract(off)
  int last = -1;
  for (int r = 0; r < m; ++r) {
    int best = 0x7fffffff, q = -1;
    for (int k = 0; k < m; ++k) {
      const int c = sl[k], idc = a.b.id[c];
      if (idc > last && idc < best) {
        best = idc;
        q = c;
      }
    }
    last = best;
    if (q < 0 || q == g) continue;
    const double reach = reach_i + a.b.R[q] + a.b.A[q];
    const double dy = min_image(yi - a.ey[q], Ly);
    const double dx = min_image(xi - a.ex[q], Lx);
    const double flo = dx + glo * dy, fhi = dx + ghi * dy;
    if (fabs(dy) < reach + s && (fabs(flo) < reach + sx || fabs(fhi) < reach + sx)) {
      if (n < kNbrCap) a.nbr[(size_t)n * a.ntot + g] = q;
      ++n;
    }
  }
}

// Build particle g's row of the list from the cells of the evaluated state.
// Cells are visited in a fixed order and each bucket's members in increasing
// particle id, so the row (and every sum over it) depends neither on the
// atomic order of the buckets nor on the storage order of the bank.  The
// sorting width is the wavefront's largest bucket, rounded up to 4, 8 or 16.
__device__ int build_row(const ListArgs& a, const PackInfo& P, double Lx, double Ly, int g,
                         double xi, double yi, double reach_i, double glo, double ghi) {
#pragma clang fp contract(off)
  double xw = xi, yw = yi;
  wrap_coord(xw, Lx);
  wrap_coord(yw, Ly);
  int col, row;
  cell_col_row(xw, yw, Lx, Ly, P, col, row);
  int rs[3], cs[3], nr = 0, nc = 0;
  for (int d = -1; d <= 1; ++d) {
    int r = (row + d + P.rows) % P.rows;
    bool dup = false;
    for (int k = 0; k < nr; ++k) dup |= rs[k] == r;
    if (!dup) rs[nr++] = r;
    int c = (col + d + P.cols) % P.cols;
    dup = false;
    for (int k = 0; k < nc; ++k) dup |= cs[k] == c;
    if (!dup) cs[nc++] = c;
  }
  // skin = the packing's cell padding; the x margin covers the largest strain
  // before the remap (selection rule above)
  const double s = P.pad, sx = fmax(1.5, 1.0 + 0.5 * Lx / Ly) * P.pad;
  int n = 0;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) {
      const int cell = P.coff + rs[i] * P.cols + cs[j];
      const int m = min(a.cnt[cell], P.cap);
      const int* sl = a.slot + (size_t)cell * P.cap;
      if (__all(m <= 4))
        row_from_bucket<4>(a, sl, m, g, xi, yi, reach_i, Lx, Ly, s, sx, glo, ghi, n);
      else if (__all(m <= 8))
        row_from_bucket<8>(a, sl, m, g, xi, yi, reach_i, Lx, Ly, s, sx, glo, ghi, n);
      else if (m <= kCapMin)
        row_from_bucket<kCapMin>(a, sl, m, g, xi, yi, reach_i, Lx, Ly, s, sx, glo, ghi, n);
      else
        row_from_crowded_bucket(a, sl, m, g, xi, yi, reach_i, Lx, Ly, s, sx, glo, ghi, n);
    }
  if (n > kNbrCap) {
    atomicOr(a.err, kErrList);
    n = kNbrCap;
  }
  a.nbr_n[g] = n;
  return n;
}

// LANES threads share a particle: lane l takes list entries l, l + LANES, ...
// in batches of kGather whose loads are all issued before any is used, so a
// lane has several independent gathers in flight.  The LANES partial sums are
// combined with xor shuffles (fixed order for a given LANES).
constexpr int kGather = 4;

// MIXED: the kernel carries the FP32-force pair path besides the FP64 one and
// picks per packing (a mixed-mode FIRE moves packings to FP64 one by one).
// FP64 systems launch the MIXED = false kernel, whose single pair path needs
// fewer registers.  Its force pass and corrector are held to 128 VGPRs (4
// wavefronts per SIMD instead of 3, at the cost of a few spilled dwords
// outside the pair loop): 1.2 us off the corrector at N = 262144 and 5% off
// the minimiser's force pass.  The predictor fits in 118 without the bound.
// SE: the plain force pass accumulates stress and energy over its contacts
// (calc_se, recorded steps).  The minimiser's passes need only the residual
// and FIRE sums, which are formed after the pair loop; their SE = false
// kernel keeps no accumulators live across the loop.
// BOX: the Heun passes change the box (resize, pure shear).  Sheared Heun
// steps run the BOX = false instance, whose epilogue carries only the
// Lees-Edwards path (corrector 114 VGPRs instead of 128 plus spills).
template <int MODE, int LANES, bool TWO_POINT, bool MIXED, bool SE, bool BOX>
__global__ void __launch_bounds__(256)
    __attribute__((amdgpu_waves_per_eu(MODE != kModePred && !MIXED ? 4 : 1))) k_list(ListArgs a) {
#pragma clang fp contract(off)
  __shared__ double s_red[4 * kNSum];
  extern __shared__ int s_hit[];  // [slot][thread] partner columns, kHitCap x blockDim
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (a.cnt_clear)
    for (int i = blockIdx.x * nthr + tid; i < a.ncells_total; i += gridDim.x * nthr)
      a.cnt_clear[i] = 0;
  if (a.flag_clear && blockIdx.x == 0)
    for (int i = tid; i < a.npack; i += nthr) a.flag_clear[i] = 0;

  const int bslot = a.block_idx ? a.block_idx[blockIdx.x] : (int)blockIdx.x;
  const int2 pb = a.pblocks[bslot];
  const int pk = pb.x;
  const PackInfo P = a.packs[pk];
  const PackScal S = a.scal[pk];
  if (MODE == kModeForces && (a.do_fire || a.do_gd) && S.converged) return;
  const int lane = tid % LANES;
  const int g = pb.y + tid / LANES;
  const bool active = g < P.poff + P.n;

  // Only the plain force pass reduces (stress/energy records, FIRE/GD sums,
  // residual); the Heun passes carry no accumulators, which keeps their
  // register footprint and occupancy where the hot loop needs it.
  constexpr bool kSums = MODE == kModeForces;
  const bool do_se = kSums && SE &&
                     (a.se_always || (a.se_every > 0 && S.step > 0 && S.step % a.se_every == 0));
  const bool do_sum = kSums && (do_se || a.do_fire || a.do_gd || a.do_resid);

  // box changes of this pass (compile-time absent in the BOX = false kernels)
  const bool box_resize = BOX && a.resize, box_pure = BOX && a.pure;
  const bool box_shrink = BOX && a.shrink, box_corrected_image = BOX && a.corrected_image;
  const double dt = P.dt, gd = P.strain_rate;
  // Lees-Edwards remap in the general form (src/cuda/find_neighbors.cu:320-330,
  // :369; src/cuda/strain_dynamics.cu:712): at gamma >= Lx/(2 Ly), x += (Lx/Ly) y
  // and gamma -= Lx/Ly.  For the square box of the CPU code the ratio is exactly
  // 1 and this is its set_back_gamma.
  const double le_ratio = S.Lx / S.Ly;
  double Lx = S.Lx, Ly = S.Ly, gam = S.gamma;
  double L_new_x = Lx, L_new_y = Ly;
  if (MODE != kModeForces && box_resize) {
    L_new_x = box_shrink ? Lx * P.resize_scale : Lx / P.resize_scale;
    L_new_y = box_shrink ? Ly * P.resize_scale : Ly / P.resize_scale;
  }
  if (MODE != kModeForces && box_pure) {
    // the box of pure_strain_step_2p (src/cuda/2point_dynamics.cu:2273-2274)
    L_new_x = Lx + 0.5 * dt * gd * Lx;
    L_new_y = Ly - 0.5 * dt * gd * Ly;
  }
  if (MODE == kModeCorr) {
    if (box_pure) {  // forces of the predicted state in the new box, strain unchanged
      Lx = L_new_x;
      Ly = L_new_y;
    } else {
      gam = S.gamma + dt * gd;
      if (box_resize && box_corrected_image) {
        Lx = L_new_x;
        Ly = L_new_y;
      }
    }
  }
  const bool fp64 = !MIXED || S.prec == kFP64;

  double acc[kSums ? kNSum : 1];
#pragma unroll
  for (int i = 0; i < (kSums ? kNSum : 1); ++i) acc[i] = 0.0;

  if (active) {
    const double xi = a.ex[g], yi = a.ey[g];
    const double nxi = a.enx[g], nyi = a.eny[g];
    const double Ri = a.b.R[g], Ai = a.b.A[g];
    const int idi = a.b.id[g];
    const double reach_i = Ri + Ai;

    // ---- list (rebuilt here when the previous pass asked for it)
    // Every lane of the particle builds the same row (identical stores), so a
    // lane only ever reads back what it wrote itself; rebuilds are rare.
    const bool rebuild = a.flag_in[pk] != 0;
    double center;
    int n;
    if (rebuild) {
      center = gam;
      n = build_row(a, P, Lx, Ly, g, xi, yi, reach_i, gam - a.gamma_window, gam + a.gamma_window);
      if (lane == 0) {
        a.xb[g] = xi;
        a.yb[g] = yi;
        if (g == P.poff) a.list_center[pk] = gam;
      }
    } else {
      n = a.nbr_n[g];
      center = a.list_center[pk];
    }

    // ---- forces over the row, in two phases.  Phase 1 walks the row with
    // batched gathers and keeps the partners that pass the bounding-circle
    // test in this thread's LDS column; phase 2 runs the closest-approach math
    // over that compacted list.  A wavefront then executes the expensive pair
    // math as often as its busiest lane has real candidates, not as often as
    // its longest row; partners are visited in row order either way, so the
    // sums are the same.
    double Fx = 0.0, Fy = 0.0, Tq = 0.0;
    int k0 = lane;  // this lane's next row entry
    while (true) {
      int nh = 0;
      for (; k0 < n && nh <= kHitCap - kGather; k0 += LANES * kGather) {
        int qs[kGather];
#pragma unroll
        for (int j = 0; j < kGather; ++j) {
          const int k = k0 + j * LANES;
          qs[j] = k < n ? a.nbr[(size_t)k * a.ntot + g] : g;
        }
        double xq[kGather], yq[kGather], Sq[kGather];
#pragma unroll
        for (int j = 0; j < kGather; ++j) {
          xq[j] = a.ex[qs[j]];
          yq[j] = a.ey[qs[j]];
          Sq[j] = a.b.R[qs[j]] + a.b.A[qs[j]];
        }
#pragma unroll
        for (int j = 0; j < kGather; ++j) {
          if (k0 + j * LANES >= n) break;
          double dx = min_image(xi - xq[j], Lx);
          double dy = min_image(yi - yq[j], Ly);
          dx += gam * dy;
          const double reach = reach_i + Sq[j];
          if (!(dx * dx + dy * dy < reach * reach * (1.0 + 1e-10))) continue;
          s_hit[nh * nthr + tid] = qs[j];
          ++nh;
        }
      }
      for (int h = 0; h < nh; ++h) {
        const int q = s_hit[h * nthr + tid];
        pair_into<kSums && SE, TWO_POINT>(xi, yi, nxi, nyi, Ri, Ai, idi, a.ex[q], a.ey[q], a.enx[q],
                                    a.eny[q], a.b.R[q], a.b.A[q], a.b.id[q], Lx, Ly, gam, fp64,
                                    do_se, Fx, Fy, Tq, acc);
      }
      if (k0 >= n) break;
    }
    if (LANES > 1) {
#pragma unroll
      for (int off = LANES >> 1; off > 0; off >>= 1) {
        Fx += __shfl_xor(Fx, off);
        Fy += __shfl_xor(Fy, off);
        Tq += __shfl_xor(Tq, off);
      }
    }
    // ---- per-particle epilogue (lane 0)
    const double half_skin = 0.5 * P.pad;
    // build position of the row (written above by this lane when it rebuilt
    // the row), loaded only here so it is not live across the pair loop
    double xbi = 0.0, ybi = 0.0;
    if (MODE != kModeForces && lane == 0) {
      xbi = a.xb[g];
      ybi = a.yb[g];
    }
    bool raise = false;
    if (lane != 0) {
    } else if (MODE == kModeForces) {
      a.b.fx[g] = Fx;
      a.b.fy[g] = Fy;
      a.b.tau[g] = Tq;
      if (kSums && do_sum) acc[kSumR2 % (kSums ? kNSum : 1)] += Fx * Fx + Fy * Fy + Tq * Tq;
      if (kSums && a.do_fire) {
        double sq = sqrt(a.b.I[g]);
        double gp = Tq / sq;
        double vx = a.b.vx[g], vy = a.b.vy[g], vp = a.b.vpsi[g];
        acc[kSumP % (kSums ? kNSum : 1)] += Fx * vx + Fy * vy + gp * vp;
        acc[kSumG % (kSums ? kNSum : 1)] += Fx * Fx + Fy * Fy + gp * gp;
        acc[kSumV % (kSums ? kNSum : 1)] += vx * vx + vy * vy + vp * vp;
      }
    } else if (MODE == kModePred) {
      const double p0 = a.b.phi[g];
      double v0x, v0y, v0t, qx, qy, qp;
      if (box_pure) {
        // pure_euler_est_2p (src/cuda/2point_dynamics.cu:1885-1890), mobility 1:
        // the affine field (+x/2, -y/2) times the rate, and the rod's rotation
        // by the strain, -RC sin(2 phi) / 2 times the rate; v0 keeps the force
        // part for the corrector
        v0x = Fx;
        v0y = Fy;
        v0t = Tq / a.b.I[g];
        qx = xi + dt * (Fx + 0.5 * gd * xi);
        qy = yi + dt * (Fy - 0.5 * gd * yi);
        qp = p0 + dt * (v0t + gd * (-0.5 * a.b.RC[g] * sin(2.0 * p0)));
      } else if (!box_resize) {
        // strain_step predictor, src/cpp/strain_dynamics.cpp:189-198
        v0x = Fx - S.gamma * Fy;
        v0y = Fy;
        v0t = Tq / a.b.I[g] - gd * iso_rot_seg(a.b.RC[g], nxi, nyi, Ai);
        qx = xi + dt * v0x;
        qy = yi + dt * v0y;
        qp = p0 + dt * v0t;
      } else {
        // resize_step predictor, src/cpp/resize_dynamics.cpp:174-190
        v0x = Fx;
        v0y = Fy;
        v0t = Tq;
        qx = xi + dt * v0x;
        qy = yi + dt * v0y;
        qp = p0 + dt * v0t / a.b.I[g];
        if (box_shrink) {
          qx *= P.resize_scale;
          qy *= P.resize_scale;
        } else {
          qx /= P.resize_scale;
          qy /= P.resize_scale;
        }
      }
      a.px[g] = qx;
      a.py[g] = qy;
      a.pphi[g] = qp;
      double sn, cs;
      sincos(qp, &sn, &cs);
      a.pnx[g] = Ai * cs;
      a.pny[g] = Ai * sn;
      a.v0x[g] = v0x;
      a.v0y[g] = v0y;
      a.v0t[g] = v0t;
      const double Lbx = (box_resize || box_pure) ? L_new_x : Lx;
      const double Lby = (box_resize || box_pure) ? L_new_y : Ly;
      double wx = qx, wy = qy;
      bool okx = wrap_coord(wx, Lbx), oky = wrap_coord(wy, Lby);
      if (!okx || !oky || !isfinite(qp)) atomicOr(a.err, kErrNonFinite);
      bin_into(g, wx, wy, Lbx, Lby, P, a.cnt_fill, a.slot_fill, a.err);
      // the corrector evaluates x~ at gamma + dt gdot, with the rescaled box
      // length in resize runs
      const double gnext = S.gamma + dt * gd;
      raise = fabs(min_image(qx - xbi, Lbx)) >= half_skin ||
              fabs(min_image(qy - ybi, Lby)) >= half_skin ||
              fabs(gnext - center) >= a.gamma_window || box_resize || box_pure;
    } else {  // kModeCorr
      double x1 = a.b.x[g], y1 = a.b.y[g], p1 = a.b.phi[g];
      double Lbx = Lx, Lby = Ly;
      bool remap = false;
      if (box_pure) {
        // pure_heun_corr_2p (src/cuda/2point_dynamics.cu:2141-2166): the
        // predicted state plus half the change of the force part
        x1 = a.px[g] + 0.5 * dt * (Fx - a.v0x[g]);
        y1 = a.py[g] + 0.5 * dt * (Fy - a.v0y[g]);
        p1 = a.pphi[g] + 0.5 * dt * (Tq / a.b.I[g] - a.v0t[g]);
        Lbx = L_new_x;
        Lby = L_new_y;
      } else if (!box_resize) {
        // strain_step corrector, src/cpp/strain_dynamics.cpp:205-215
        double v1x = Fx - gam * Fy;
        double v1y = Fy;
        // the isolated rotation at the predicted angle: (nxi, nyi) is the
        // predicted segment vector, the state this pass evaluates
        double v1t = Tq / a.b.I[g] - gd * iso_rot_seg(a.b.RC[g], nxi, nyi, Ai);
        x1 += 0.5 * dt * (a.v0x[g] + v1x);
        y1 += 0.5 * dt * (a.v0y[g] + v1y);
        p1 += 0.5 * dt * (a.v0t[g] + v1t);
        double gnew = S.gamma + gd * dt;
        if (gnew >= 0.5 * le_ratio) {  // set_back_gamma, src/cpp/spherocyl_box.cpp:529-540
          remap = true;
          if (y1 < 0) y1 += S.Ly;
          else if (y1 > S.Ly) y1 -= S.Ly;
          x1 += le_ratio * y1;
          if (x1 < 0) x1 += S.Lx;
          else {
            int guard = 0;
            while (x1 > S.Lx && guard++ < 8) x1 -= S.Lx;
          }
        }
        Lbx = S.Lx;
        Lby = S.Ly;
      } else {
        // resize_step corrector, src/cpp/resize_dynamics.cpp:196-215
        x1 += 0.5 * dt * (a.v0x[g] + Fx);
        y1 += 0.5 * dt * (a.v0y[g] + Fy);
        p1 += 0.5 * dt * (a.v0t[g] + Tq) / a.b.I[g];
        if (box_shrink) {
          x1 *= P.resize_scale;
          y1 *= P.resize_scale;
        } else {
          x1 /= P.resize_scale;
          y1 /= P.resize_scale;
        }
        Lbx = L_new_x;
        Lby = L_new_y;
      }
      bool okx = wrap_coord(x1, Lbx), oky = wrap_coord(y1, Lby);
      if (!okx || !oky || !isfinite(p1)) atomicOr(a.err, kErrNonFinite);
      a.b.x[g] = x1;
      a.b.y[g] = y1;
      a.b.phi[g] = p1;
      double sn, cs;
      sincos(p1, &sn, &cs);
      a.b.nx[g] = Ai * cs;
      a.b.ny[g] = Ai * sn;
      bin_into(g, x1, y1, Lbx, Lby, P, a.cnt_fill, a.slot_fill, a.err);
      double gnext = box_pure ? S.gamma : S.gamma + gd * dt;
      if (remap) gnext -= le_ratio;
      raise = remap || box_resize || box_pure || fabs(min_image(x1 - xbi, Lbx)) >= half_skin ||
              fabs(min_image(y1 - ybi, Lby)) >= half_skin || fabs(gnext - center) >= a.gamma_window;
    }
    if (raise && a.flag_out) a.flag_out[pk] = 1;
  }

  // ---- scalar advance (corrector: one thread per packing)
  if (MODE == kModeCorr && g == P.poff && lane == 0) {
    // copied word by word and updated in place: the struct never lives in
    // registers across the kernel (it would spill)
    const int* src = reinterpret_cast<const int*>(&a.scal[pk]);
    int* dst = reinterpret_cast<int*>(&a.scal_next[pk]);
    for (int w = 0; w < (int)(sizeof(PackScal) / sizeof(int)); ++w) dst[w] = src[w];
    PackScal& S2 = a.scal_next[pk];
    S2.step = S2.step + 1;
    if (box_pure) {
      S2.Lx = L_new_x;
      S2.Ly = L_new_y;
      S2.gamma_total = S2.gamma_total + dt * gd;
    } else if (!box_resize) {
      const double ratio = S2.Lx / S2.Ly;
      double g2 = S2.gamma + gd * dt, gt = S2.gamma_total + gd * dt;
      if (g2 >= 0.5 * ratio) {
        g2 -= ratio;
        if (ratio == 1.0) gt = (double)(int)(gt + 1) + g2;
      }
      S2.gamma = g2;
      S2.gamma_total = gt;
    } else {
      S2.Lx = L_new_x;
      S2.Ly = L_new_y;
    }
  }

  if constexpr (kSums) {
    if (!do_sum) return;
    block_sum<kNSum>(acc, s_red, nthr);
    if (tid < kNSum) a.part[(size_t)bslot * kNSum + tid] = s_red[tid];
  }
}

#define SC_LIST_INST(L, T, X)                                                    \
  template __global__ void k_list<kModeForces, L, T, X, true, false>(ListArgs);  \
  template __global__ void k_list<kModeForces, L, T, X, false, false>(ListArgs); \
  template __global__ void k_list<kModePred, L, T, X, false, false>(ListArgs);   \
  template __global__ void k_list<kModeCorr, L, T, X, false, false>(ListArgs);   \
  template __global__ void k_list<kModePred, L, T, X, false, true>(ListArgs);    \
  template __global__ void k_list<kModeCorr, L, T, X, false, true>(ListArgs);
#define SC_LIST_INST_LANES(T, X) \
  SC_LIST_INST(1, T, X)          \
  SC_LIST_INST(2, T, X)          \
  SC_LIST_INST(4, T, X)          \
  SC_LIST_INST(8, T, X)
SC_LIST_INST_LANES(false, false)
SC_LIST_INST_LANES(true, false)
SC_LIST_INST_LANES(false, true)
SC_LIST_INST_LANES(true, true)
#undef SC_LIST_INST_LANES
#undef SC_LIST_INST

// (nx, ny) = A (cos phi, sin phi) for every particle of the bank.
__global__ void k_segvec(Bank b, int ntot) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  double sn, cs;
  sincos(b.phi[g], &sn, &cs);
  b.nx[g] = b.A[g] * cs;
  b.ny[g] = b.A[g] * sn;
}

}  // namespace sc
