// Device kernels for gfx950 (MI355X).
//
// Work decomposition
//   * Particles of each packing are stored sorted by cell (counting sort, see
//     k_scan_cells / k_sort_cells / k_gather).  Cell membership is kept exact
//     every step in fixed-capacity buckets (cnt[c], slot[c*kCap+k]) that the
//     position-update kernels refill with one atomic per particle, so no
//     neighbour-list rebuild trigger and no host round trip is needed.
//   * A workgroup owns a tile: a run of cells of one cell row.  It stages the
//     particles of the 3 x (width+2) cells around the tile into LDS (position,
//     segment vector A(cos phi, sin phi), R, A, id), ordered by particle id within
//     each cell so every sum below has a fixed order.
//   * L lanes cooperate on one own particle.  Phase A tests all candidates of the
//     9 surrounding cells with a cheap centre-distance bound and compacts the
//     survivors into a per-particle LDS list (ballot + popcount, order kept);
//     phase B runs the segment-segment closest-approach force on the list, so
//     the expensive math is spent on real near pairs only.  The L partial sums
//     are combined with xor shuffles.
//   * Every pair is evaluated in the reference's orientation (lower particle id
//     first, src/cpp/spherocyl_box.cpp:514-520), so both particles of a pair see
//     bit-identical pair results.
//   * Energy / stress / FIRE sums reduce per tile, then the last tile of each
//     packing (ticket counter) reduces the tile partials in tile order.
#include <hip/hip_runtime.h>

#include "sc_internal.h"
#include "sc_physics.h"

namespace sc {

// LDS carve-up of the tile kernel (sizes in entries).
struct TileLds {
  int* cs;             // staged cell prefix [128]
  double *x, *y, *nx, *ny, *R, *A;
  int *id, *g;
  unsigned char* col;
  unsigned short* list;
  double* red;
};

__device__ __forceinline__ TileLds carve(unsigned char* base, int cap, int groups) {
  TileLds t;
  size_t off = 0;
  t.cs = reinterpret_cast<int*>(base + off);
  off += 128 * sizeof(int);
  t.x = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.y = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.nx = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.ny = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.R = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.A = reinterpret_cast<double*>(base + off);
  off += (size_t)cap * 8;
  t.id = reinterpret_cast<int*>(base + off);
  off += (size_t)cap * 4;
  t.g = reinterpret_cast<int*>(base + off);
  off += (size_t)cap * 4;
  t.col = base + off;
  off += ((size_t)cap + 15) / 16 * 16;
  t.list = reinterpret_cast<unsigned short*>(base + off);
  off += (size_t)groups * kListLen * 2;
  off = (off + 15) / 16 * 16;
  t.red = reinterpret_cast<double*>(base + off);
  return t;
}

// Host mirror of carve(): bytes of dynamic LDS for a launch.
size_t tile_lds_bytes(int cap, int groups, int nthr) {
  size_t off = 128 * sizeof(int) + (size_t)cap * (6 * 8 + 4 + 4);
  off += ((size_t)cap + 15) / 16 * 16;
  off += (size_t)groups * kListLen * 2;
  off = (off + 15) / 16 * 16;
  off += (size_t)(nthr / 64) * kNSum * 8;
  return off;
}

// Minimiser decision for a packing from its reduced sums: the FIRE step
// (velocity mixing or retreat, time step and mixing factor, convergence) or
// the gradient-descent stop test.  S is the state the force pass evaluated.
__device__ PackScal decide(const ReduceArgs& a, const PackInfo& P, const PackScal& S,
                           const double* tot) {
#pragma clang fp contract(off)
  PackScal S2 = S;
  const double frms = sqrt(tot[kSumR2] / (3.0 * P.n));
  if (a.do_fire) {
    const FireParams& fp = a.fire;
    S2.frms = frms;
    // Convergence is judged on FP64 pair forces only: a packing still on FP32
    // pair math whose residual falls below max(switch_frms, ftol) moves to FP64
    // and keeps iterating.
    const bool fp32 = S.prec == kFP32;
    if (!fp32 && frms < fp.ftol) {
      S2.converged = 1;
      S2.iters = S.step;
    } else if (S.step >= (int)fp.max_iter) {
      S2.converged = 2;
      S2.iters = S.step;
    } else {
      double P_ = tot[kSumP], G_ = tot[kSumG], V_ = tot[kSumV];
      double dt = S.fire_dt, alpha = S.fire_alpha, npos = S.fire_npos;
      S2.fire_dt_old = dt;
      if (P_ >= 0) {
        S2.fire_mix = 1;
        S2.fire_c_keep = 1.0 - alpha;
        S2.fire_c_force = (G_ > 0) ? alpha * sqrt(V_ / G_) : 0.0;
        npos += 1;
        if (npos > fp.n_min) {
          dt = fmin(dt * fp.f_inc, fp.dt_max);
          alpha = alpha * fp.f_alpha;
        }
      } else {
        S2.fire_mix = 0;
        S2.fire_c_keep = 0;
        S2.fire_c_force = 0;
        dt = dt * fp.f_dec;
        alpha = fp.alpha0;
        npos = 0;
      }
      S2.fire_dt = dt;
      S2.fire_dt_new = dt;
      S2.fire_alpha = alpha;
      S2.fire_npos = npos;
      S2.step = S.step + 1;
      if (fp32 && frms < fmax(fp.switch_frms, fp.ftol)) S2.prec = kFP64;
    }
  } else if (a.do_gd) {
    // gradient-descent passes always record energy: E = tot[kSeE] / (2N)
    S2.frms = frms;
    if (tot[kSeE] * P.inv_2n <= a.fire.emin || S.step >= (int)a.fire.max_iter) {
      S2.converged = 1;
      S2.iters = S.step;
    } else {
      S2.step = S.step + 1;
    }
  }
  return S2;
}

// Finalisation of the reduced sums of packing pk (one thread): stress/energy
// records, and the minimiser decision into scal_next.
__device__ void finalize_pack(const ReduceArgs& a, int pk, const PackInfo& P, const PackScal& S,
                              const double* tot, bool do_se) {
#pragma clang fp contract(off)
  double* ps = a.pack_sum + (size_t)pk * kNSum;
  double area = S.Lx * S.Ly;
  if (do_se) {
    ps[kSeE] = tot[kSeE] * P.inv_2n;
    for (int i = kSePxx; i <= kSeCPyx; ++i) ps[i] = tot[i] / area;
    ps[kSeNc] = tot[kSeNc];
    if (a.se_ring && a.se_every > 0 && S.step > 0 && S.step % a.se_every == 0) {
      int rec = S.step / a.se_every - 1;
      double* r = a.se_ring + ((size_t)rec * a.npack + pk) * kNSe;
      for (int i = 0; i < kNSe; ++i) r[i] = ps[i];
    }
  }
  ps[kSumP] = tot[kSumP];
  ps[kSumG] = tot[kSumG];
  ps[kSumV] = tot[kSumV];
  ps[kSumR2] = tot[kSumR2];
  if (a.do_fire || a.do_gd) a.scal_next[pk] = decide(a, P, S, tot);
}

// Sum packing pk's workgroup partials in workgroup order: thread t takes
// partials t, t + nthr, ..., then the fixed-order block sum.  Every thread
// calls it; the totals are in s_red[0 .. kNSum) afterwards.
__device__ void reduce_partials(const ReduceArgs& a, int pk, double* s_red) {
#pragma clang fp contract(off)
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int2 rg = a.ranges[pk];
  double tot[kNSum];
#pragma unroll
  for (int i = 0; i < kNSum; ++i) tot[i] = 0.0;
  for (int t = rg.x + tid; t < rg.x + rg.y; t += nthr) {
    const double* tp = a.part + (size_t)t * kNSum;
#pragma unroll
    for (int i = 0; i < kNSum; ++i) tot[i] += tp[i];
  }
  block_sum<kNSum>(tot, s_red, nthr);
}

__device__ bool records_se(const ReduceArgs& a, const PackScal& S) {
  return a.se_always || (a.se_every > 0 && S.step > 0 && S.step % a.se_every == 0);
}

// One workgroup per packing: reduce, then finalise.  Runs as its own launch
// after the force pass, so the partials are visible without device-scope
// fences in the force kernel.  (Minimisers over packings of few workgroups
// reduce inside k_update instead.)
__global__ void __launch_bounds__(256) k_reduce(ReduceArgs a) {
  __shared__ double s_red[4 * kNSum];
  const int pk = a.pack_idx ? a.pack_idx[blockIdx.x] : (int)blockIdx.x;
  const PackInfo P = a.packs[pk];
  const PackScal S = a.scal[pk];
  if ((a.do_fire || a.do_gd) && S.converged) {
    if (threadIdx.x == 0) a.scal_next[pk] = S;
    return;
  }
  reduce_partials(a, pk, s_red);
  if (threadIdx.x == 0) {
    double fin[kNSum];
    for (int i = 0; i < kNSum; ++i) fin[i] = s_red[i];
    finalize_pack(a, pk, P, S, fin, records_se(a, S));
  }
}

// ---------------------------------------------------------------------------
// the tile force kernel
// ---------------------------------------------------------------------------
template <int MODE, bool TWO_POINT>
__global__ void __launch_bounds__(256) k_tile(TileArgs a) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;

  if (a.cnt_clear)
    for (int i = blockIdx.x * nthr + tid; i < a.ncells_total; i += gridDim.x * nthr)
      a.cnt_clear[i] = 0;

  const int4 tl = a.tiles[blockIdx.x];
  const int pk = tl.x, row = tl.y, c0 = tl.z, ntw = tl.w;
  const PackInfo P = a.packs[pk];
  const PackScal S = a.scal[pk];

  if (MODE == kModeForces && (a.do_fire || a.do_gd) && S.converged) return;

  const bool do_se =
      (MODE != kModeCorr) &&
      (a.se_always || (a.se_every > 0 && S.step > 0 && S.step % a.se_every == 0));
  const bool do_sum = do_se || (MODE == kModeForces && (a.do_fire || a.do_gd || a.do_resid));

  TileLds L = carve(lds_raw, a.stage_cap, a.list_groups);

  // ---- the staged 3 x (ntw + 2) block of cells
  const int rows = P.rows, cols = P.cols;
  int nr, ci, r0, r1, r2;
  if (rows >= 3) {
    nr = 3;
    ci = 1;
    r0 = (row + rows - 1) % rows;
    r1 = row;
    r2 = (row + 1) % rows;
  } else if (rows == 2) {
    nr = 2;
    ci = 0;
    r0 = row;
    r1 = 1 - row;
    r2 = row;
  } else {
    nr = 1;
    ci = 0;
    r0 = row;
    r1 = row;
    r2 = row;
  }
  const bool full = cols < ntw + 2;  // narrow box: the whole cell row is staged once
  const int nc = full ? cols : ntw + 2;
  const int ncell = nr * nc;

  for (int k = tid; k < ncell; k += nthr) {
    int rr = k / nc, jj = k - rr * nc;
    int r = rr == 0 ? r0 : (rr == 1 ? r1 : r2);
    int cc = full ? jj : (c0 - 1 + jj + cols) % cols;
    int cell = P.coff + r * cols + cc;
    L.cs[k + 1] = min(a.cnt[cell], P.cap);
  }
  __syncthreads();
  if (tid < 64) {  // exclusive prefix of <= 128 cell counts in wave 0
    int i0 = 2 * tid + 1, i1 = 2 * tid + 2;
    int v0 = i0 <= ncell ? L.cs[i0] : 0;
    int v1 = i1 <= ncell ? L.cs[i1] : 0;
    int sum = v0 + v1;
    for (int off = 1; off < 64; off <<= 1) {
      int t = __shfl_up(sum, off);
      if (tid >= off) sum += t;
    }
    int excl = sum - v0 - v1;
    if (i0 <= ncell) L.cs[i0] = excl + v0;
    if (i1 <= ncell) L.cs[i1] = excl + v0 + v1;
    if (tid == 0) L.cs[0] = 0;
  }
  __syncthreads();
  const int total = L.cs[ncell];

  // ---- stage particles, ordered by id inside each cell
  for (int e = tid; e < total; e += nthr) {
    int lo = 0, hi = ncell;  // find k: cs[k] <= e < cs[k+1]
    while (hi - lo > 1) {
      int mid = (lo + hi) >> 1;
      if (L.cs[mid] <= e) lo = mid;
      else hi = mid;
    }
    const int k = lo;
    int rr = k / nc, jj = k - rr * nc;
    int r = rr == 0 ? r0 : (rr == 1 ? r1 : r2);
    int cc = full ? jj : (c0 - 1 + jj + cols) % cols;
    int cell = P.coff + r * cols + cc;
    int m = L.cs[k + 1] - L.cs[k];
    int j = e - L.cs[k];
    const int* sl = a.slot + (size_t)cell * P.cap;
    int g = sl[j];
    int myid = a.b.id[g];
    int rank = 0;
    for (int q = 0; q < m; ++q) rank += (a.b.id[sl[q]] < myid);
    int dst = L.cs[k] + rank;
    double ph = a.sphi[g];
    double Ai = a.b.A[g];
    double sn, cs;
    sincos(ph, &sn, &cs);
    L.x[dst] = a.sx[g];
    L.y[dst] = a.sy[g];
    L.nx[dst] = Ai * cs;
    L.ny[dst] = Ai * sn;
    L.R[dst] = a.b.R[g];
    L.A[dst] = Ai;
    L.id[dst] = myid;
    L.g[dst] = g;
    L.col[dst] = (unsigned char)jj;
  }
  __syncthreads();

  const int own_lo = L.cs[ci * nc + (full ? 0 : 1)];
  const int own_hi = L.cs[ci * nc + (full ? nc : 1 + ntw)];

  // box scalars for this pass
  const double dt = P.dt, gd = P.strain_rate;
  double Lx = S.Lx, Ly = S.Ly;
  double gam = S.gamma;
  double L_new_x = Lx, L_new_y = Ly;
  if (MODE != kModeForces && a.resize) {
    L_new_x = a.shrink ? Lx * P.resize_scale : Lx / P.resize_scale;
    L_new_y = a.shrink ? Ly * P.resize_scale : Ly / P.resize_scale;
  }
  if (MODE == kModeCorr) {
    gam = S.gamma + dt * gd;
    if (a.resize && a.corrected_image) {
      Lx = L_new_x;
      Ly = L_new_y;
    }
  }
  const bool fp64 = S.prec == kFP64;

  // kLanes lanes per particle, fixed so that every per-particle sum has the same
  // order whatever the tile shape or block size
  constexpr int LN = kLanes;
  const int G = nthr / LN;
  const int grp = tid / LN, lane = tid - grp * LN;
  const int grp_shift = (tid & 63) - lane;
  const unsigned long long gmask = (1ull << LN) - 1ull;
  unsigned short* my_list = L.list + grp * kListLen;

  double acc[kNSum];
#pragma unroll
  for (int i = 0; i < kNSum; ++i) acc[i] = 0.0;

  for (int base_o = own_lo; base_o < own_hi; base_o += G) {
    const int o = base_o + grp;
    if (o >= own_hi) continue;
    const double xi = L.x[o], yi = L.y[o], Ri = L.R[o], Ai = L.A[o];
    const int idi = L.id[o];
    const int jc = L.col[o];
    int lo_[3], n_[3];
#pragma unroll
    for (int rr = 0; rr < 3; ++rr) {
      if (rr < nr) {
        int k0 = rr * nc + (full ? 0 : jc - 1), k1 = rr * nc + (full ? nc : jc + 2);
        lo_[rr] = L.cs[k0];
        n_[rr] = L.cs[k1] - lo_[rr];
      } else {
        lo_[rr] = 0;
        n_[rr] = 0;
      }
    }
    const int ntot = n_[0] + n_[1] + n_[2];
    const double reach_i = Ri + Ai;

    // phase A: centre-distance bound, compaction into my_list (order kept)
    int nlist = 0;
    for (int base = 0; base < ntot; base += LN) {
      const int k = base + lane;
      bool pass = false;
      int e = 0;
      if (k < ntot) {
        e = k < n_[0] ? lo_[0] + k : (k < n_[0] + n_[1] ? lo_[1] + (k - n_[0]) : lo_[2] + (k - n_[0] - n_[1]));
        if (L.id[e] != idi) {
          double dx = min_image(xi - L.x[e], Lx);
          double dy = min_image(yi - L.y[e], Ly);
          dx += gam * dy;
          double reach = reach_i + L.R[e] + L.A[e];
          pass = dx * dx + dy * dy < reach * reach * (1.0 + 1e-10);
        }
      }
      unsigned long long m = (__ballot(pass) >> grp_shift) & gmask;
      if (pass) {
        int pos = nlist + __popcll(m & ((1ull << lane) - 1ull));
        if (pos < kListLen) my_list[pos] = (unsigned short)e;
      }
      nlist += __popcll(m);
    }
    if (nlist > kListLen) {
      if (lane == 0) atomicOr(a.err, kErrList);
      nlist = kListLen;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // phase B: closest-approach force on the compacted list
    double Fx = 0.0, Fy = 0.0, Tq = 0.0;
    for (int k = lane; k < nlist; k += LN) {
      const int e = my_list[k];
      const bool iP = idi < L.id[e];
      const int p = iP ? o : e, q = iP ? e : o;
      double ddx = min_image(L.x[p] - L.x[q], Lx);
      double ddy = min_image(L.y[p] - L.y[q], Ly);
      ddx += gam * ddy;
      const double sig = L.R[p] + L.R[q];
      PairOut<double> r0, r1;
      bool h0, h1;
      pair_contacts<TWO_POINT>(ddx, ddy, L.nx[p], L.ny[p], L.nx[q], L.ny[q], L.A[p], L.A[q], sig,
                               fp64, r0, h0, r1, h1);
      if (h0) {
        if (iP) {
          Fx += r0.fx;
          Fy += r0.fy;
          Tq += r0.tp;
          if (do_se) {
            add_contact_sums(r0, L.nx[p], L.ny[p], L.nx[q], L.ny[q], sig, acc);
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
          if (do_se) add_contact_sums(r1, L.nx[p], L.ny[p], L.nx[q], L.ny[q], sig, acc);
        } else {
          Fx -= r1.fx;
          Fy -= r1.fy;
          Tq += r1.tq;
        }
      }
    }
    for (int off = LN >> 1; off > 0; off >>= 1) {
      Fx += __shfl_xor(Fx, off);
      Fy += __shfl_xor(Fy, off);
      Tq += __shfl_xor(Tq, off);
    }
    if (lane != 0) continue;

    // ---- per-particle epilogue
    const int g = L.g[o];
    if (MODE == kModeForces) {
      a.b.fx[g] = Fx;
      a.b.fy[g] = Fy;
      a.b.tau[g] = Tq;
      if (do_sum) acc[kSumR2] += Fx * Fx + Fy * Fy + Tq * Tq;
      if (a.do_fire) {
        double sq = sqrt(a.b.I[g]);
        double gp = Tq / sq;
        double vx = a.b.vx[g], vy = a.b.vy[g], vp = a.b.vpsi[g];
        acc[kSumP] += Fx * vx + Fy * vy + gp * vp;
        acc[kSumG] += Fx * Fx + Fy * Fy + gp * gp;
        acc[kSumV] += vx * vx + vy * vy + vp * vp;
      }
    } else if (MODE == kModePred) {
      const double x0 = xi, y0 = yi, p0 = a.b.phi[g];
      double v0x, v0y, v0t, qx, qy, qp;
      if (!a.resize) {
        // strain_step predictor, src/cpp/strain_dynamics.cpp:189-198
        v0x = Fx - S.gamma * Fy;
        v0y = Fy;
        v0t = Tq / a.b.I[g] - gd * iso_rot_seg(a.b.RC[g], L.nx[o], L.ny[o], L.A[o]);
        qx = x0 + dt * v0x;
        qy = y0 + dt * v0y;
        qp = p0 + dt * v0t;
      } else {
        // resize_step predictor, src/cpp/resize_dynamics.cpp:174-190
        v0x = Fx;
        v0y = Fy;
        v0t = Tq;
        qx = x0 + dt * v0x;
        qy = y0 + dt * v0y;
        qp = p0 + dt * v0t / a.b.I[g];
        if (a.shrink) {
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
      a.v0x[g] = v0x;
      a.v0y[g] = v0y;
      a.v0t[g] = v0t;
      // the corrector's force pass runs on the cells of the predicted state
      double Lbx = a.resize ? L_new_x : Lx, Lby = a.resize ? L_new_y : Ly;
      double wx = qx, wy = qy;
      bool okx = wrap_coord(wx, Lbx), oky = wrap_coord(wy, Lby);
      if (!okx || !oky || !isfinite(qp)) atomicOr(a.err, kErrNonFinite);
      bin_into(g, wx, wy, Lbx, Lby, P, a.cnt_fill, a.slot_fill, a.err);
      if (do_sum) acc[kSumR2] += Fx * Fx + Fy * Fy + Tq * Tq;
    } else {  // kModeCorr
      double x1 = a.b.x[g], y1 = a.b.y[g], p1 = a.b.phi[g];
      double Lbx = Lx, Lby = Ly;
      if (!a.resize) {
        // strain_step corrector, src/cpp/strain_dynamics.cpp:205-215
        double v1x = Fx - gam * Fy;
        double v1y = Fy;
        double v1t = Tq / a.b.I[g] - gd * iso_rot_seg(a.b.RC[g], L.nx[o], L.ny[o], L.A[o]);
        double ddx = 0.5 * dt * (a.v0x[g] + v1x);
        double ddy = 0.5 * dt * (a.v0y[g] + v1y);
        double ddp = 0.5 * dt * (a.v0t[g] + v1t);
        x1 += ddx;
        y1 += ddy;
        p1 += ddp;
        double gnew = S.gamma + gd * dt;
        const double ratio = S.Lx / S.Ly;  // general Lees-Edwards remap, as k_list
        if (gnew >= 0.5 * ratio) {  // set_back_gamma, src/cpp/spherocyl_box.cpp:529-540
          if (y1 < 0) y1 += S.Ly;
          else if (y1 > S.Ly) y1 -= S.Ly;
          x1 += ratio * y1;
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
        double ddx = 0.5 * dt * (a.v0x[g] + Fx);
        double ddy = 0.5 * dt * (a.v0y[g] + Fy);
        double ddp = 0.5 * dt * (a.v0t[g] + Tq) / a.b.I[g];
        x1 += ddx;
        y1 += ddy;
        p1 += ddp;
        if (a.shrink) {
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
      bin_into(g, x1, y1, Lbx, Lby, P, a.cnt_fill, a.slot_fill, a.err);
    }
  }

  // ---- scalar advance of this packing (corrector: one thread per packing)
  if (MODE == kModeCorr && blockIdx.x == P.toff && tid == 0) {
    PackScal S2 = S;
    S2.step = S.step + 1;
    if (!a.resize) {
      const double ratio = S.Lx / S.Ly;
      S2.gamma = S.gamma + gd * dt;
      S2.gamma_total = S.gamma_total + gd * dt;
      if (S2.gamma >= 0.5 * ratio) {
        S2.gamma -= ratio;
        if (ratio == 1.0) S2.gamma_total = (double)(int)(S2.gamma_total + 1) + S2.gamma;
      }
    } else {
      S2.Lx = L_new_x;
      S2.Ly = L_new_y;
    }
    a.scal_next[pk] = S2;
  }

  if (!do_sum) return;
  block_sum<kNSum>(acc, L.red, nthr);
  if (tid < kNSum) a.tile_part[(size_t)blockIdx.x * kNSum + tid] = L.red[tid];
}

template __global__ void k_tile<kModeForces, false>(TileArgs);
template __global__ void k_tile<kModePred, false>(TileArgs);
template __global__ void k_tile<kModeCorr, false>(TileArgs);
template __global__ void k_tile<kModeForces, true>(TileArgs);
template __global__ void k_tile<kModePred, true>(TileArgs);
template __global__ void k_tile<kModeCorr, true>(TileArgs);

// ---------------------------------------------------------------------------
// particle-parallel kernels
// ---------------------------------------------------------------------------

// Wrap every particle (set_particle / find_cells rules) and bin it.
__global__ void k_bin_all(Bank b, const int* pack_of, const PackInfo* packs, const PackScal* scal,
                          int* cnt, int* slot, int* err, int ntot, int wrap_angles) {
#pragma clang fp contract(off)
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  int pk = pack_of[g];
  const PackInfo P = packs[pk];
  const PackScal S = scal[pk];
  double x = b.x[g], y = b.y[g];
  bool okx = wrap_coord(x, S.Lx), oky = wrap_coord(y, S.Ly);
  bool ok = okx && oky;
  if (!ok) atomicOr(err, kErrNonFinite);
  b.x[g] = x;
  b.y[g] = y;
  if (wrap_angles) {
    double p = b.phi[g];
    wrap_angle(p);
    b.phi[g] = p;
  }
  bin_into(g, x, y, S.Lx, S.Ly, P, cnt, slot, err);
}

// Gradient-descent / FIRE step after a force pass (one thread per particle,
// each workgroup within one packing) + binning.  With u.fused every workgroup
// first reduces its packing's partials and takes the minimiser decision itself
// (the same sums in the same order in every workgroup, so the same decision);
// the packing's first workgroup records it.  That saves a launch per
// iteration, and the previous kernel's partials are visible without fences.
// Packings of many force workgroups are reduced once by k_reduce instead
// (u.fused = 0), whose decision is read here.
// mode 0: gradient_relax_step (src/cpp/relaxation_dynamics.cpp:15-19)
// mode 1: FIRE update (specification in oracle/sc_oracle.cpp)
// With neighbour lists (flag_out != null) the update raises the packing's
// rebuild flag when a particle leaves its skin/2 box around the build position.
__global__ void __launch_bounds__(256) k_update(UpdateArgs u) {
#pragma clang fp contract(off)
  __shared__ double s_red[4 * kNSum];
  __shared__ PackScal s_dec;
  const Bank& b = u.b;
  const int tid = threadIdx.x;
  const int lin = blockIdx.x * blockDim.x + tid;
  if (u.flag_clear)
    for (int i = lin; i < u.npack; i += gridDim.x * blockDim.x) u.flag_clear[i] = 0;
  // this workgroup's packing and its decision (uniform over the workgroup)
  const int2 ub = u.ublocks[blockIdx.x];
  const int pk = ub.x;
  const PackInfo P = u.packs[pk];
  // The packing's workgroups clear its cells of the bucket buffer the next
  // pass fills (workgroup w of nw takes cells w*256 + tid, stepping nw*256).
  // Packings left out of a compacted launch keep their buckets: they have
  // finished, and every bucket is refilled before the state is evaluated again.
  if (u.cnt_clear) {
    const int w = (ub.y - P.poff) >> 8, nw = (P.n + 255) >> 8, ncell = P.rows * P.cols;
    for (int c = w * 256 + tid; c < ncell; c += nw * 256) u.cnt_clear[P.coff + c] = 0;
  }
  const PackScal S0 = u.red.scal[pk];
  const bool lead = ub.y == P.poff;  // the packing's first workgroup stores the decision
  // A finished packing is left as it is: no move, no store, no binning.  Its
  // force passes return at once, and the buckets of every packing are filled
  // again before the state is evaluated after the minimiser.
  if (S0.converged) {
    if (u.fused && lead && tid == 0) u.red.scal_next[pk] = S0;
    return;
  }
  // The particle's state is loaded before the reduction, so these loads are in
  // flight while the partials are summed.
  const int g = ub.y + tid;
  const bool own = g < P.poff + P.n;
  const int gl = own ? g : P.poff;
  double x = b.x[gl], y = b.y[gl], p = b.phi[gl];
  const double Fx = b.fx[gl], Fy = b.fy[gl], Tq = b.tau[gl], Ig = b.I[gl];
  const double Ag = b.A[gl];
  double vx = 0.0, vy = 0.0, vp = 0.0;
  if (u.mode != 0) {
    vx = b.vx[gl];
    vy = b.vy[gl];
    vp = b.vpsi[gl];
  }
  double xbg = 0.0, ybg = 0.0;
  if (u.flag_out) {
    xbg = u.xb[gl];
    ybg = u.yb[gl];
  }
  if (u.fused) {
    reduce_partials(u.red, pk, s_red);
    if (tid == 0) {
      double fin[kNSum];
      for (int i = 0; i < kNSum; ++i) fin[i] = s_red[i];
      if (lead) finalize_pack(u.red, pk, P, S0, fin, records_se(u.red, S0));
      s_dec = decide(u.red, P, S0, fin);
    }
  } else if (tid == 0) {
    s_dec = u.red.scal_next[pk];
  }
  __syncthreads();
  const PackScal S = s_dec;
  if (S.converged) return;  // converged at the state just evaluated
  int bin_c = -1;
  if (own) {
    if (u.mode == 0) {
      double dt = P.dt;
      x += dt * Fx;
      y += dt * Fy;
      p += dt * Tq / Ig;
    } else {
      double sq = sqrt(Ig);
      double gx = Fx, gy = Fy, gp = Tq / sq;
      double ux, uy, up;
      if (S.fire_mix) {
        ux = S.fire_c_keep * vx + S.fire_c_force * gx;
        uy = S.fire_c_keep * vy + S.fire_c_force * gy;
        up = S.fire_c_keep * vp + S.fire_c_force * gp;
      } else {
        x -= 0.5 * S.fire_dt_old * vx;
        y -= 0.5 * S.fire_dt_old * vy;
        p -= 0.5 * S.fire_dt_old * vp / sq;
        ux = uy = up = 0.0;
      }
      double dtn = S.fire_dt_new;
      ux += dtn * gx;
      uy += dtn * gy;
      up += dtn * gp;
      b.vx[g] = ux;
      b.vy[g] = uy;
      b.vpsi[g] = up;
      double mx = dtn * ux, my = dtn * uy;
      double m2 = mx * mx + my * my;
      if (m2 > u.max_move * u.max_move) {
        double sc = u.max_move / sqrt(m2);
        mx *= sc;
        my *= sc;
      }
      x += mx;
      y += my;
      p += dtn * up / sq;
    }
    bool okx = wrap_coord(x, S.Lx), oky = wrap_coord(y, S.Ly);
    if (!okx || !oky || !isfinite(p)) atomicOr(u.err, kErrNonFinite);
    b.x[g] = x;
    b.y[g] = y;
    b.phi[g] = p;
    double sn, cs;
    sincos(p, &sn, &cs);
    b.nx[g] = Ag * cs;
    b.ny[g] = Ag * sn;
    bin_c = cell_of(x, y, S.Lx, S.Ly, P);
    if (u.flag_out) {
      const double hs = 0.5 * P.pad;  // list skin = the packing's cell padding
      if (fabs(min_image(x - xbg, S.Lx)) >= hs || fabs(min_image(y - ybg, S.Ly)) >= hs)
        u.flag_out[pk] = 1;
    }
  }
  bin_wave<1>(bin_c, g, P.cap, u.cnt, u.slot, u.err);
}

// FIRE of one small packing per workgroup, from its current state to
// convergence (or max_iter), in one launch.  An iteration is the force pass,
// the reduction and decision, and the update of the two-launch path (k_list
// force pass, k_update), with the same formulas and the same FIRE state
// machine (decide); only the order of the sums differs.  Nothing of an
// iteration touches global memory: the packing's coordinates, segment vectors,
// radii, half lengths, ids, rows and forces live in LDS, and each particle's
// angle, velocity and build position in the registers of the thread that
// updates it.
//
// Two mappings of threads to particles:
//   * force pass and row builds: L lanes per particle (the largest power of
//     two up to 8 with n L <= kSmallThreads), lane l taking row entries
//     l, l + L, ...; the lanes' sums are combined with xor shuffles and lane 0
//     stores the particle's force and torque in LDS;
//   * update: thread t moves particle t, so the FIRE sums, the velocity mixing
//     and the move (square roots, divisions, sincos) run once per particle, on
//     ceil(n / 64) wavefronts, not once per lane.
// A row (the partners within the skin: the selection rule of the list kernels
// at the fixed strain of a minimisation) is rebuilt from all n particles when
// a particle has moved half the skin from its build position, as the update
// kernel decides for the global rows; a ballot places the hits in increasing
// partner index.
template <bool TWO_POINT, bool MIXED>
__global__ void __launch_bounds__(kSmallThreads) k_fire_small(SmallFireArgs a) {
#pragma clang fp contract(off)
  constexpr int kRowStride = kSmallRow + 2;  // odd word stride: rows start in different banks
  constexpr int kWaves = kSmallThreads / 64;
  constexpr int kFire = 4;                   // FIRE sums: |F|^2 + T^2, F.v, |g|^2, |v|^2
  __shared__ double s_x[kSmallMax], s_y[kSmallMax], s_nx[kSmallMax], s_ny[kSmallMax];
  __shared__ double s_R[kSmallMax], s_A[kSmallMax];
  __shared__ double s_fx[kSmallMax], s_fy[kSmallMax], s_tq[kSmallMax];
  __shared__ int s_id[kSmallMax];
  __shared__ unsigned short s_row[kSmallMax * kRowStride];
  __shared__ double s_part[kWaves * kFire];
  __shared__ PackScal s_S;
  // Flags, each written and read in phases separated by barriers so that
  // every thread sees the same value and all leave the loop together:
  // s_rebuild (raised by the update, read at the top of the next iteration,
  // cleared during the decision), s_stop (non-finite coordinate: raised by the
  // update, read at the top), s_over (row overflow: raised by a rebuild, read
  // after its barrier).
  __shared__ int s_rebuild, s_stop, s_over;
  const Bank& b = a.b;
  const int tid = threadIdx.x, wave = tid >> 6;
  const int pk = a.pack_idx[blockIdx.x];
  const PackInfo P = a.packs[pk];
  const int n = P.n;
  int L = 1, lg = 0;
  while (L < 8 && n * 2 * L <= kSmallThreads) {
    L *= 2;
    ++lg;
  }
  // force mapping
  const int lane = tid & (L - 1);
  const int p = tid >> lg;
  const bool own = p < n;
  const int glane = (tid & 63) & ~(L - 1);  // first lane of the particle's group in the wave
  // update mapping
  const bool upd = tid < n;
  const int nuw = (n + 63) >> 6;  // wavefronts holding update threads
  const int gu = P.poff + (upd ? tid : 0);
  double x = b.x[gu], y = b.y[gu], phi = b.phi[gu];
  double vx = b.vx[gu], vy = b.vy[gu], vp = b.vpsi[gu];
  const double sq = sqrt(b.I[gu]), Au = b.A[gu];
  if (upd) {
    s_x[tid] = x;
    s_y[tid] = y;
    s_nx[tid] = b.nx[gu];
    s_ny[tid] = b.ny[gu];
    s_R[tid] = b.R[gu];
    s_A[tid] = Au;
    s_id[tid] = b.id[gu];
  }
  if (tid == 0) {
    s_S = a.scal[pk];
    s_rebuild = 1;
    s_stop = 0;
    s_over = 0;
  }
  __syncthreads();
  PackScal S = s_S;
  const double Lx = S.Lx, Ly = S.Ly, gam = S.gamma;
  // the list kernels' selection rule and skin (the x margin covers the
  // largest strain before the remap, Lx / (2 Ly))
  const double skin = P.pad, sx = fmax(1.5, 1.0 + 0.5 * Lx / Ly) * P.pad, hs = 0.5 * P.pad;
  const FireParams& fp = a.fire;
  double xb = x, yb = y;
  int cnt = 0;  // row length of the force mapping's particle (same in its L lanes)
  for (int steps = 0;; ++steps) {
    if (s_stop || steps == a.max_steps) break;
    if (s_rebuild) {
      xb = x;
      yb = y;
      cnt = 0;
      const int pp = own ? p : 0;
      const double xp = s_x[pp], yp = s_y[pp], reach_p = s_R[pp] + s_A[pp];
      for (int base = 0; base < n; base += L) {
        const int q = base + lane;
        bool hit = false;
        if (own && q < n && q != p) {
          const double reach = reach_p + s_R[q] + s_A[q];
          const double dy = min_image(yp - s_y[q], Ly);
          const double dx = min_image(xp - s_x[q], Lx);
          const double f = dx + gam * dy;
          hit = fabs(dy) < reach + skin && fabs(f) < reach + sx;
        }
        const unsigned long long bits = (__ballot(hit) >> glane) & ((1ull << L) - 1ull);
        const int pos = cnt + __popcll(bits & ((1ull << lane) - 1ull));
        if (hit && pos < kSmallRow) s_row[p * kRowStride + pos] = (unsigned short)q;
        cnt += __popcll(bits);
      }
      if (cnt > kSmallRow) {
        if (lane == 0) atomicOr(a.err, kErrList);
        s_over = 1;
      }
      __syncthreads();  // rows and s_over visible
      if (s_over) break;
    }

    // ---- force pass (force mapping): bounding-circle hits of this lane's
    // entries as a bit mask, then the pair math over the set bits
    const bool fp64 = !MIXED || S.prec == kFP64;
    double Fx = 0.0, Fy = 0.0, Tq = 0.0;
    if (own) {
      const double xp = s_x[p], yp = s_y[p], nxp = s_nx[p], nyp = s_ny[p];
      const double Rp = s_R[p], Ap = s_A[p], reach_p = Rp + Ap;
      const int idp = s_id[p];
      const unsigned short* row = s_row + p * kRowStride;
      unsigned long long mask = 0;
      for (int k = lane, j = 0; k < cnt; k += L, ++j) {
        const int q = row[k];
        double dx = min_image(xp - s_x[q], Lx);
        const double dy = min_image(yp - s_y[q], Ly);
        dx += gam * dy;
        const double reach = reach_p + s_R[q] + s_A[q];
        if (dx * dx + dy * dy < reach * reach * (1.0 + 1e-10)) mask |= 1ull << j;
      }
      while (mask) {
        const int j = __ffsll((long long)mask) - 1;
        mask &= mask - 1;
        const int q = row[lane + j * L];
        pair_into<false, TWO_POINT>(xp, yp, nxp, nyp, Rp, Ap, idp, s_x[q], s_y[q], s_nx[q],
                                    s_ny[q], s_R[q], s_A[q], s_id[q], Lx, Ly, gam, fp64, false,
                                    Fx, Fy, Tq, nullptr);
      }
    }
    for (int off = L >> 1; off > 0; off >>= 1) {
      Fx += __shfl_xor(Fx, off);
      Fy += __shfl_xor(Fy, off);
      Tq += __shfl_xor(Tq, off);
    }
    if (own && lane == 0) {
      s_fx[p] = Fx;
      s_fy[p] = Fy;
      s_tq[p] = Tq;
    }
    __syncthreads();  // forces visible to the update mapping

    // ---- FIRE sums (update mapping), per wavefront in lane order
    double fx = 0.0, fy = 0.0, tq = 0.0, gp = 0.0;
    if (wave < nuw) {
      double acc[kFire] = {0.0, 0.0, 0.0, 0.0};
      if (upd) {
        fx = s_fx[tid];
        fy = s_fy[tid];
        tq = s_tq[tid];
        gp = tq / sq;
        acc[0] = fx * fx + fy * fy + tq * tq;
        acc[1] = fx * vx + fy * vy + gp * vp;
        acc[2] = fx * fx + fy * fy + gp * gp;
        acc[3] = vx * vx + vy * vy + vp * vp;
      }
#pragma unroll
      for (int i = 0; i < kFire; ++i) {
        double v = acc[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        acc[i] = v;
      }
      if ((tid & 63) == 0)
#pragma unroll
        for (int i = 0; i < kFire; ++i) s_part[wave * kFire + i] = acc[i];
    }
    __syncthreads();  // partials visible
    if (tid == 0) {
      double tot[kNSum];
      for (int i = 0; i < kNSum; ++i) tot[i] = 0.0;
      for (int w = 0; w < nuw; ++w) {
        tot[kSumR2] += s_part[w * kFire + 0];
        tot[kSumP] += s_part[w * kFire + 1];
        tot[kSumG] += s_part[w * kFire + 2];
        tot[kSumV] += s_part[w * kFire + 3];
      }
      ReduceArgs ra{};
      ra.fire = fp;
      ra.do_fire = 1;
      s_S = decide(ra, P, S, tot);
      s_rebuild = 0;  // every thread read it before the barriers above
    }
    __syncthreads();  // decision visible
    S = s_S;
    if (S.converged) break;  // converged (or max_iter) at the state just evaluated

    // ---- update (update mapping; the FIRE branch of k_update)
    if (upd) {
      double ux, uy, up;
      if (S.fire_mix) {
        ux = S.fire_c_keep * vx + S.fire_c_force * fx;
        uy = S.fire_c_keep * vy + S.fire_c_force * fy;
        up = S.fire_c_keep * vp + S.fire_c_force * gp;
      } else {
        x -= 0.5 * S.fire_dt_old * vx;
        y -= 0.5 * S.fire_dt_old * vy;
        phi -= 0.5 * S.fire_dt_old * vp / sq;
        ux = uy = up = 0.0;
      }
      const double dtn = S.fire_dt_new;
      ux += dtn * fx;
      uy += dtn * fy;
      up += dtn * gp;
      vx = ux;
      vy = uy;
      vp = up;
      double mx = dtn * ux, my = dtn * uy;
      const double m2 = mx * mx + my * my;
      if (m2 > fp.max_move * fp.max_move) {
        const double sc = fp.max_move / sqrt(m2);
        mx *= sc;
        my *= sc;
      }
      x += mx;
      y += my;
      phi += dtn * up / sq;
      const bool okx = wrap_coord(x, Lx), oky = wrap_coord(y, Ly);
      if (!okx || !oky || !isfinite(phi)) {
        atomicOr(a.err, kErrNonFinite);
        s_stop = 1;
      }
      double sn, cs;
      sincos(phi, &sn, &cs);
      s_x[tid] = x;
      s_y[tid] = y;
      s_nx[tid] = Au * cs;
      s_ny[tid] = Au * sn;
      if (fabs(min_image(x - xb, Lx)) >= hs || fabs(min_image(y - yb, Ly)) >= hs) s_rebuild = 1;
    }
    __syncthreads();  // new coordinates, s_rebuild and s_stop visible
  }
  if (upd) {
    b.x[gu] = x;
    b.y[gu] = y;
    b.phi[gu] = phi;
    b.nx[gu] = s_nx[tid];
    b.ny[gu] = s_ny[tid];
    b.vx[gu] = vx;
    b.vy[gu] = vy;
    b.vpsi[gu] = vp;
    b.fx[gu] = s_fx[tid];
    b.fy[gu] = s_fy[tid];
    b.tau[gu] = s_tq[tid];
  }
  if (tid == 0) a.scal[pk] = S;
}
template __global__ void k_fire_small<false, false>(SmallFireArgs);
template __global__ void k_fire_small<true, false>(SmallFireArgs);
template __global__ void k_fire_small<false, true>(SmallFireArgs);
template __global__ void k_fire_small<true, true>(SmallFireArgs);

// Affine rescale (x, y) *= scale[pack] (quasi-static compression step).
__global__ void k_affine(Bank b, const int* pack_of, const double* scale, int ntot, int y_only) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  double s = scale[pack_of[g]];
  if (!y_only) b.x[g] *= s;
  b.y[g] *= s;
}

// ---- counting sort by cell ------------------------------------------------

// Exclusive scan of bucket counts over the cells of each packing (one
// workgroup per packing, chunks of blockDim cells).
__global__ void k_scan_cells(const PackInfo* packs, const int* cnt, int* cstart) {
  __shared__ int s_w[16];
  __shared__ int s_carry;
  const PackInfo P = packs[blockIdx.x];
  const int ncell = P.rows * P.cols;
  const int tid = threadIdx.x, nthr = blockDim.x;
  if (tid == 0) s_carry = P.poff;
  __syncthreads();
  for (int base = 0; base < ncell; base += nthr) {
    int c = base + tid;
    int v = c < ncell ? min(cnt[P.coff + c], P.cap) : 0;
    int incl = v;
    for (int off = 1; off < 64; off <<= 1) {
      int t = __shfl_up(incl, off);
      if ((tid & 63) >= off) incl += t;
    }
    if ((tid & 63) == 63) s_w[tid >> 6] = incl;
    __syncthreads();
    int wpre = 0;
    for (int w = 0; w < (tid >> 6); ++w) wpre += s_w[w];
    int carry = s_carry;
    if (c < ncell) cstart[P.coff + c] = carry + wpre + incl - v;
    __syncthreads();
    if (tid == nthr - 1) s_carry = carry + wpre + incl;
    __syncthreads();
  }
}

// Per cell: order members by particle id, record the permutation and rewrite
// the bucket with the new (sorted) storage indices.
__global__ void k_sort_cells(const int* cstart, int* cnt, int* slot, const int* id, int* perm,
                             int ncells_total, int cap) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells_total) return;
  int m = min(cnt[c], cap);
  int* sl = slot + (size_t)c * cap;
  const int base = cstart[c];
  if (m <= kCapMin) {
    int mem[kCapMin], key[kCapMin];
    for (int j = 0; j < m; ++j) {
      mem[j] = sl[j];
      key[j] = id[mem[j]];
    }
    for (int j = 1; j < m; ++j) {  // insertion sort by id
      int km = key[j], mm = mem[j], i = j - 1;
      while (i >= 0 && key[i] > km) {
        key[i + 1] = key[i];
        mem[i + 1] = mem[i];
        --i;
      }
      key[i + 1] = km;
      mem[i + 1] = mm;
    }
    for (int j = 0; j < m; ++j) perm[base + j] = mem[j];
  } else {
    // a crowded cell: insertion sort by id in the bucket itself
    for (int j = 1; j < m; ++j) {
      const int mm = sl[j], km = id[mm];
      int i = j - 1;
      while (i >= 0 && id[sl[i]] > km) {
        sl[i + 1] = sl[i];
        --i;
      }
      sl[i + 1] = mm;
    }
    for (int j = 0; j < m; ++j) perm[base + j] = sl[j];
  }
  for (int j = 0; j < m; ++j) sl[j] = base + j;
  cnt[c] = m;
}

// Coalesced gather into the other bank: dst[d] = src[perm[d]].
__global__ void k_gather(Bank src, Bank dst, const int* perm, int ntot, int with_velocity) {
  int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ntot) return;
  int s = perm[d];
  // after a bucket overflow the permutation has holes (the overflow is reported
  // to the host); keep every read in bounds regardless
  if (s < 0 || s >= ntot) s = d;
  dst.x[d] = src.x[s];
  dst.y[d] = src.y[s];
  dst.phi[d] = src.phi[s];
  dst.R[d] = src.R[s];
  dst.A[d] = src.A[s];
  dst.I[d] = src.I[s];
  dst.RC[d] = src.RC[s];
  dst.fx[d] = src.fx[s];
  dst.fy[d] = src.fy[s];
  dst.tau[d] = src.tau[s];
  dst.id[d] = src.id[s];
  dst.nx[d] = src.nx[s];
  dst.ny[d] = src.ny[s];
  if (with_velocity) {
    dst.vx[d] = src.vx[s];
    dst.vy[d] = src.vy[s];
    dst.vpsi[d] = src.vpsi[s];
  }
}

// ---- neighbour / contact export (diagnostics) ------------------------------

// For every particle of packing pk: the reference's neighbour criterion
// (check_neighbors, src/cpp/spherocyl_box.cpp:303-318) or the contact test over
// the own + 8 adjacent cells (deduplicated), partners with a larger id only.
// Pairs go to a per-particle region of width kExport.
constexpr int kExport = 64;
__global__ void k_export(Bank b, const PackInfo* packs, const PackScal* scal, const int* cnt,
                         const int* slot, int pk, int contacts, int* out_q, int* out_n) {
#pragma clang fp contract(off)
  const PackInfo P = packs[pk];
  const PackScal S = scal[pk];
  int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= P.n) return;
  int g = P.poff + li;
  int col, row;
  cell_col_row(b.x[g], b.y[g], S.Lx, S.Ly, P, col, row);
  int rs[3], cs[3], nr = 0, ncl = 0;
  for (int d = -1; d <= 1; ++d) {
    int r = (row + d + P.rows) % P.rows;
    bool dup = false;
    for (int k = 0; k < nr; ++k) dup |= rs[k] == r;
    if (!dup) rs[nr++] = r;
    int c = (col + d + P.cols) % P.cols;
    dup = false;
    for (int k = 0; k < ncl; ++k) dup |= cs[k] == c;
    if (!dup) cs[ncl++] = c;
  }
  int idp = b.id[g];
  int count = 0;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < ncl; ++j) {
      int cell = P.coff + rs[i] * P.cols + cs[j];
      int m = min(cnt[cell], P.cap);
      for (int k = 0; k < m; ++k) {
        int q = slot[(size_t)cell * P.cap + k];
        int idq = b.id[q];
        if (idq <= idp) continue;
        bool hit;
        if (!contacts) {
          double reach = b.R[g] + b.A[g] + b.R[q] + b.A[q];
          double dy = min_image(b.y[g] - b.y[q], S.Ly);
          hit = false;
          if (fabs(dy) < reach + P.pad) {
            double dx = min_image(b.x[g] - b.x[q], S.Lx);
            double dxp = dx + (0.5 * (S.Lx / S.Ly)) * dy;  // remap strain 0.5 Lx/Ly
            dx += S.gamma * dy;
            hit = fabs(dx) < reach + P.pad || fabs(dxp) < reach + P.pad;
          }
        } else {
          double ddx = min_image(b.x[g] - b.x[q], S.Lx);
          double ddy = min_image(b.y[g] - b.y[q], S.Ly);
          ddx += S.gamma * ddy;
          PairOut<double> r;
          hit = pair_contact_f64(ddx, ddy, b.A[g] * cos(b.phi[g]), b.A[g] * sin(b.phi[g]),
                                 b.A[q] * cos(b.phi[q]), b.A[q] * sin(b.phi[q]), b.A[g], b.A[q],
                                 b.R[g] + b.R[q], r);
        }
        if (hit) {
          if (count < kExport) out_q[(size_t)li * kExport + count] = idq;
          ++count;
        }
      }
    }
  out_n[li] = count;
}

// ---- caller-order transfer -------------------------------------------------

// out[poff + id[g]] = field(g) for the selected fields (positions or forces).
__global__ void k_to_caller(Bank b, const int* pack_of, const PackInfo* packs, int what,
                            double* o0, double* o1, double* o2, int ntot) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  int dst = packs[pack_of[g]].poff + b.id[g];
  if (what == 0) {
    o0[dst] = b.x[g];
    o1[dst] = b.y[g];
    o2[dst] = b.phi[g];
  } else {
    o0[dst] = b.fx[g];
    o1[dst] = b.fy[g];
    o2[dst] = b.tau[g];
  }
}

// field(g) = in[poff + id[g]] for positions.
__global__ void k_from_caller(Bank b, const int* pack_of, const PackInfo* packs, const double* i0,
                              const double* i1, const double* i2, int ntot) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ntot) return;
  int src = packs[pack_of[g]].poff + b.id[g];
  b.x[g] = i0[src];
  b.y[g] = i1[src];
  b.phi[g] = i2[src];
}

}  // namespace sc
