// Internal definitions shared by the device kernels (sc_device.hip) and the host
// side of the C ABI (sc_api.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sc {

// Particles a cell bucket holds: at least kCapMin (the width of the in-register
// sorts of a bucket), raised by the host to twice the largest occupancy it
// sees (polydisperse packings, whose cells are sized by the longest rod, put
// many short rods in one cell), at most kCapMax.  The capacity of a system is
// PackInfo::cap, the same for all its packings.
constexpr int kCapMin = 16;
constexpr int kCapMax = 256;
// Columns of cells per tile (upper bound; the autotuner picks the actual width).
constexpr int kMaxTileCols = 32;
// Passing candidates kept per particle in the LDS contact list.
constexpr int kListLen = 48;
// Lanes that cooperate on one particle (fixed: it sets the summation order).
constexpr int kLanes = 8;
// Particles processed concurrently per workgroup (LDS list slots).
constexpr int kMaxGroups = 256 / kLanes;
// Reduction slots: 9 stress/energy sums, then FIRE's P G V and the residual R2.
constexpr int kSeE = 0, kSePxx = 1, kSePyy = 2, kSePxy = 3, kSeCPxx = 4, kSeCPyy = 5,
              kSeCPxy = 6, kSeCPyx = 7, kSeNc = 8, kSumP = 9, kSumG = 10, kSumV = 11,
              kSumR2 = 12;
constexpr int kNSum = 13;
constexpr int kNSe = 9;

// The reference's truncated pi (src/cpp/spherocyl_box.cpp:12).
constexpr double kRefPi = 3.14159265358979;

enum Prec : int { kFP64 = 0, kFP32 = 1 };

// Static description of one packing.
struct PackInfo {
  int n, poff;           // particles poff .. poff+n-1
  int rows, cols, coff;  // cell grid, first global cell index
  int toff, ntiles;      // tile range
  int tile_cols;
  int cap;               // bucket capacity (slots per cell)
  double pad;            // cell padding
  double dt, strain_rate, resize_rate;
  double resize_scale;   // (1 - resize_rate)^dt, the box scale of one resize step
  double inv_2n;         // 1 / (2N), energy normalisation
};

// Mutable per-packing scalars, double-buffered: a kernel that advances a
// packing reads scal[cur] and one designated thread writes scal[next].
struct PackScal {
  double Lx, Ly, gamma, gamma_total;
  double fire_dt, fire_alpha, fire_npos;
  double fire_c_keep, fire_c_force, fire_dt_new, fire_dt_old;
  double frms;           // residual of the last reduced force pass
  int step;              // sub-steps / iterations since the counter was reset
  int fire_mix;          // 1: keep-and-mix velocities, 0: retreat and zero them
  int converged;         // minimiser finished (FIRE: F_rms < ftol, GD: E <= Emin)
  int prec;              // Prec of the pair math for this packing
  int iters;             // minimiser iterations at convergence
  int pad_;
};

struct FireParams {
  double dt0, dt_max, n_min, f_inc, f_dec, alpha0, f_alpha, max_move, ftol, max_iter;
  double switch_frms;    // mixed mode: FP32 pair math while frms >= switch_frms
  double emin;           // gradient descent: stop when E <= emin
};

enum ErrBit : int { kErrCellOverflow = 1, kErrList = 4, kErrNonFinite = 8 };

// Per-particle arrays of one storage bank (sorted by cell within a packing).
// (nx, ny) = A (cos phi, sin phi) is kept in step with phi by every kernel
// that moves particles, so force passes never recompute it.
struct Bank {
  double *x, *y, *phi, *R, *A, *I, *RC, *fx, *fy, *tau, *vx, *vy, *vpsi, *nx, *ny;
  int* id;
};

enum TileMode : int { kModeForces = 0, kModePred = 1, kModeCorr = 2 };

// Arguments of the tile force kernel.
struct TileArgs {
  const PackInfo* packs;
  const PackScal* scal;   // scalars of the state being evaluated
  PackScal* scal_next;    // advanced scalars (pred/corr/fire)
  const int4* tiles;      // (packing, row, first column, columns)
  const int* cnt;         // bucket counts / slots of the staged state's cells
  const int* slot;
  int* cnt_fill;          // buckets filled with the new positions (pred: x~, corr: x')
  int* slot_fill;
  int* cnt_clear;         // bucket counts zeroed by this launch (or null)
  int ncells_total;
  Bank b;                 // current bank (own state, static particle data)
  const double* sx;       // staged positions (current or predicted)
  const double* sy;
  const double* sphi;
  double *px, *py, *pphi; // predicted positions (pred writes, corr reads)
  double *v0x, *v0y, *v0t;
  double* tile_part;      // [tiles][kNSum] per-tile partial sums
  int* err;
  FireParams fire;
  int se_every;           // pred/forces: record when step % se_every == 0 (0: never)
  int se_always;          // forces: reduce stress/energy unconditionally
  int do_fire;            // forces: FIRE sums + decisions in the last block
  int do_gd;              // forces: gradient-descent stop test in the last block
  int do_resid;           // forces: reduce the residual sum(F^2) for F_rms
  int write_forces;
  int resize;             // pred/corr: resize dynamics instead of shear
  int shrink;
  int corrected_image;
  int npack;
  int stage_cap;          // LDS staging capacity (entries)
  int list_groups;        // LDS contact-list slots
  int two_point;          // two-point contact model instead of the closest point
};

// Neighbour-list force path.  Each particle keeps a list of candidate partners
// chosen with a skin; a force kernel rebuilds the lists of a packing (each
// thread its own row) when the previous kernel raised the packing's flag.
constexpr int kNbrCap = 64;
// Partners a list-kernel thread holds between its prefilter and pair-math
// phases (a longer run is processed in pieces); the kernel's dynamic LDS is
// kHitCap ints per thread.
constexpr int kHitCap = 16;

struct ListArgs {
  const PackInfo* packs;
  const PackScal* scal;
  PackScal* scal_next;
  const int2* pblocks;     // per workgroup: packing, first particle
  Bank b;                  // current bank
  const double *ex, *ey, *enx, *eny;  // evaluated state (current or predicted)
  double *px, *py, *pphi, *pnx, *pny; // predicted state (pred writes, corr reads)
  double *v0x, *v0y, *v0t;
  const int* cnt;          // cells of the evaluated state
  const int* slot;
  int* cnt_fill;           // cells of the state this pass produces
  int* slot_fill;
  int* cnt_clear;
  int ncells_total;
  int* nbr;                // [kNbrCap][ntot] partner indices
  int* nbr_n;              // [ntot]
  double *xb, *yb;         // positions at the last list build
  double* list_center;     // [npack] strain at the last list build
  const int* flag_in;      // [npack] rebuild before evaluating
  int* flag_out;           // [npack] raised for the next force pass
  int* flag_clear;         // [npack] zeroed here (written by the pass after next)
  double gamma_window;     // lists cover strains center +- gamma_window
  double* part;            // [workgroups][kNSum] partial sums
  int* err;
  int ntot, npack;
  int se_every, se_always, do_fire, do_gd, do_resid;
  int resize, shrink, corrected_image;
  int pure;                // pure shear: x expands, y contracts at the packing's rate / 2
  int two_point;           // two-point contact model instead of the closest point
  const int* block_idx;    // launched workgroup -> workgroup of pblocks (null: identity)
};

// Arguments of the per-packing reduction / finalisation kernel.
struct ReduceArgs {
  const PackInfo* packs;
  const PackScal* scal;   // scalars of the state the force pass evaluated
  PackScal* scal_next;    // FIRE / GD decisions
  const double* part;     // block partial sums [blocks][kNSum]
  const int2* ranges;     // per packing: first block, number of blocks
  double* pack_sum;       // [npack][kNSum] latest reduced (scaled) sums
  double* se_ring;        // recorded stress/energy [rec][npack][kNSe]
  FireParams fire;
  int se_every, se_always, do_fire, do_gd;
  int npack;
  const int* pack_idx;    // launched workgroup -> packing (null: identity)
};

// Arguments of the gradient-descent / FIRE update kernel.
struct UpdateArgs {
  Bank b;
  const int* pack_of;
  const PackInfo* packs;
  int* cnt;               // buckets filled with the new positions
  int* slot;
  int* cnt_clear;         // buckets cleared for the pass after next (each packing's own cells)
  int* err;
  const double *xb, *yb;  // list build positions (list path)
  int* flag_out;          // list path: raise the rebuild flag
  int* flag_clear;
  double skin, max_move;   // skin unused: the packing's pad is the skin
  int ntot, npack, mode;
  const int2* ublocks;     // per workgroup: packing, first particle (runs of <= 256)
  ReduceArgs red;          // the force pass's partials, scalars and minimiser settings
  int fused;               // 1: reduce and decide here; 0: k_reduce already wrote red.scal_next
};

// FIRE of small packings in one launch: one workgroup of kSmallThreads per
// packing of at most kSmallMax particles holds the packing in LDS and runs
// every iteration (force pass, reduction, decision, update, list rebuild)
// itself, so an iteration costs no launch and no global round trip.
constexpr int kSmallMax = 256;
constexpr int kSmallThreads = 512;
constexpr int kSmallRow = kNbrCap;  // partners a particle's LDS row holds
struct SmallFireArgs {
  Bank b;
  const PackInfo* packs;
  PackScal* scal;          // read at the start, written at the end (in place)
  const int* pack_idx;     // workgroup -> packing
  int* err;
  FireParams fire;
  int max_steps;           // iterations per launch: the host launches again for packings left
};

// Device view of a system's current state, for other device code of this
// library (the message-passing warm start).  Valid until the next call that
// moves particles or re-sorts the bank.
struct SystemView {
  int device;
  hipStream_t stream;
  int ntot, npack, ncells;
  Bank b;                  // particle arrays in storage (cell-sorted) order; b.id = caller index
  const PackInfo* packs;
  const PackScal* scal;    // current per-packing scalars
  const int* cnt;          // exact cell buckets of the current state
  const int* slot;
  const int* pack_of;      // storage index -> packing
  const double* pack_sum;  // [npack][kNSum] reduced sums of the last force pass
};

}  // namespace sc

struct sc_system;
namespace sc {
int system_view(sc_system* s, SystemView* v);
// After device code changed x, y or phi of the current bank: wrap, re-bin,
// refresh segment vectors, re-sort and schedule a list rebuild.
int system_positions_changed(sc_system* s);
// After device code changed x, y or phi of the current bank: bin, refresh
// segment vectors and evaluate energy and stress of every packing (into the
// view's pack_sum), without re-sorting the bank.  Errors (a crowded cell, a
// non-finite coordinate) surface at the next synchronising call.
int system_energy_pass(sc_system* s);
}  // namespace sc

