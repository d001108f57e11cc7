// Host side of the C ABI (include/spherocyls.h): system lifecycle, cell grids,
// tiles and workgroup maps, kernel orchestration on the system's own HIP
// stream, the neighbour-list rebuild-flag sequence, timing and the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spherocyls.h"
#include "sc_internal.h"
#include "sc_status.h"

namespace sc {
// kernels (sc_device.hip, sc_list.hip)
template <int MODE, bool TWO_POINT>
__global__ void k_tile(TileArgs a);
template <int MODE, int LANES, bool TWO_POINT, bool MIXED, bool SE, bool BOX>
__global__ void k_list(ListArgs a);
__global__ void k_reduce(ReduceArgs a);
__global__ void k_segvec(Bank b, int ntot);
__global__ void k_bin_all(Bank b, const int* pack_of, const PackInfo* packs, const PackScal* scal,
                          int* cnt, int* slot, int* err, int ntot, int wrap_angles);
__global__ void k_update(UpdateArgs u);
template <bool TWO_POINT, bool MIXED>
__global__ void k_fire_small(SmallFireArgs a);
__global__ void k_affine(Bank b, const int* pack_of, const double* scale, int ntot, int y_only);
__global__ void k_scan_cells(const PackInfo* packs, const int* cnt, int* cstart);
__global__ void k_sort_cells(const int* cstart, int* cnt, int* slot, const int* id, int* perm,
                             int ncells_total, int cap);
__global__ void k_gather(Bank src, Bank dst, const int* perm, int ntot, int with_velocity);
__global__ void k_export(Bank b, const PackInfo* packs, const PackScal* scal, const int* cnt,
                         const int* slot, int pk, int contacts, int* out_q, int* out_n);
__global__ void k_to_caller(Bank b, const int* pack_of, const PackInfo* packs, int what,
                            double* o0, double* o1, double* o2, int ntot);
__global__ void k_from_caller(Bank b, const int* pack_of, const PackInfo* packs, const double* i0,
                              const double* i1, const double* i2, int ntot);
size_t tile_lds_bytes(int cap, int groups, int nthr);
constexpr int kExport = 64;
}  // namespace sc

using namespace sc;

namespace {

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIPCK(call)                                  \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) throw HipError{e_, #call}; \
  } while (0)

// Cell grid of configure_cells (src/cpp/spherocyl_box.cpp:228-241).  Columns
// are 2.24 (Rmax + Amax) + pad wide: a pair in contact is at most
// 2 (Rmax + Amax) sqrt(1 + gamma^2) apart in sheared x, and |gamma| <= 1/2 in
// a square box.  A box with Lx != Ly remaps at |gamma| = Lx / (2 Ly), so its
// columns are widened to 2 sqrt(1 + (Lx / 2 Ly)^2) (Rmax + Amax) + pad when
// that is larger (Lx / Ly > 1.0087); square boxes keep the reference's rule.
double cell_width_factor(double Lx, double Ly) {
  const double r = Lx / Ly;
  return std::max(2.24, 2.0 * std::sqrt(1.0 + 0.25 * r * r));
}

void cell_grid(double Lx, double Ly, double Rmax, double Amax, double& pad, int& rows, int& cols) {
  double wmin = cell_width_factor(Lx, Ly) * (Rmax + Amax) + pad;
  double hmin = 2 * (Rmax + Amax) + pad;
  rows = std::max(static_cast<int>(Ly / hmin), 1);
  cols = std::max(static_cast<int>(Lx / wmin), 1);
  if (pad == 0) pad = std::max(Lx / cols - wmin, Ly / rows - hmin);
}

struct EventPair {
  hipEvent_t a, b;
  int kernel;
};

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b on XCD
// b % 8).  Reorder a packing's spatially ordered work items so that each XCD
// receives a contiguous run: neighbouring items share particles, and the
// shared data then stays in one XCD's L2.
template <typename T>
std::vector<T> xcd_order(const std::vector<T>& spatial, int first_launch_index) {
  const int nx = 8, T_ = (int)spatial.size();
  std::vector<int> count(nx, 0), start(nx, 0), used(nx, 0);
  for (int i = 0; i < T_; ++i) count[(first_launch_index + i) % nx]++;
  for (int x = 1; x < nx; ++x) start[x] = start[x - 1] + count[x - 1];
  std::vector<T> out(T_);
  for (int i = 0; i < T_; ++i) {
    int x = (first_launch_index + i) % nx;
    out[i] = spatial[start[x] + used[x]++];
  }
  return out;
}

}  // namespace

struct sc_system {
  int device = 0;
  hipStream_t stream = nullptr;
  int npack = 0, ntot = 0;
  std::vector<int> offsets;
  std::vector<PackInfo> packs;
  std::vector<PackScal> scal;  // host copy of d_scal[cur_scal] (valid after download_scal)
  std::vector<double> Rmax, Amax, filled;
  int path = SC_PATH_LIST;
  int tile_cols = 12, block = 128, list_block = 256, list_lanes = 1;
  int cap = kCapMin;  // bucket capacity (slots per cell) of every packing
  int lds_max = 64 * 1024;
  int prec = SC_PREC_FP64;
  int contact_model = SC_CONTACT_CLOSEST;
  bool use_graphs = true;
  int small_max = kSmallMax;  // FIRE of packings up to this size in one launch (0: never)
  int small_steps = 50000;    // ... for at most this many iterations per launch
  int* d_small_idx = nullptr;
  int ncells = 0, ntiles = 0, npblocks = 0;
  std::vector<int4> tiles;
  std::vector<int2> pblocks, tile_ranges, list_ranges;

  PackInfo* d_packs = nullptr;
  PackScal* d_scal[2] = {nullptr, nullptr};
  int cur_scal = 0;
  int4* d_tiles = nullptr;
  int2* d_pblocks = nullptr;
  int2* d_tile_ranges = nullptr;
  int2* d_list_ranges = nullptr;
  // Launch lists restricted to the packings a minimiser still advances
  // (list path): force workgroups (indices into pblocks), packings, and
  // update workgroups.  Used while `compact` is set.
  bool compact = false;
  int* d_act_blocks = nullptr;
  int* d_act_packs = nullptr;
  int2* d_act_ublocks = nullptr;
  int n_act_blocks = 0, n_act_packs = 0, n_act_ublocks = 0;
  // update workgroups of every packing: (packing, first particle), runs of <= 256
  std::vector<int2> ublocks;
  int2* d_ublocks = nullptr;
  int* d_cnt[3] = {nullptr, nullptr, nullptr};
  int* d_slot[3] = {nullptr, nullptr, nullptr};
  int cur_buf = 0;
  int* d_cstart = nullptr;
  int* d_perm = nullptr;
  Bank bank[2]{};
  int cur_bank = 0;
  double *d_px = nullptr, *d_py = nullptr, *d_pphi = nullptr, *d_pnx = nullptr, *d_pny = nullptr;
  double *d_v0x = nullptr, *d_v0y = nullptr, *d_v0t = nullptr;
  int* d_pack_of = nullptr;
  double* d_tile_part = nullptr;
  double* d_list_part = nullptr;
  double* d_pack_sum = nullptr;
  double* d_se_ring = nullptr;
  size_t se_ring_cap = 0;
  int* d_err = nullptr;
  double* d_io = nullptr;  // 3 * ntot staging for caller-order transfers
  int* d_export_q = nullptr;
  int* d_export_n = nullptr;
  int export_cap = 0;
  double* d_scale = nullptr;
  bool velocity_live = false;
  // neighbour lists
  int* d_nbr = nullptr;
  int* d_nbr_n = nullptr;
  double *d_xb = nullptr, *d_yb = nullptr;
  double* d_list_center = nullptr;
  int* d_flag[3] = {nullptr, nullptr, nullptr};
  long long kseq = 0;  // position in the force/update kernel sequence
  double gamma_window = 0.05;

  bool timing = false;
  std::vector<EventPair> ev_pending;
  std::vector<hipEvent_t> ev_free;
  double kt_ms[SC_KERNEL_COUNT] = {};
  long long kt_n[SC_KERNEL_COUNT] = {};

  std::string err;
  std::vector<void*> allocs;

  template <typename T>
  T* dalloc(size_t count) {
    void* p = nullptr;
    HIPCK(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
    allocs.push_back(p);
    return static_cast<T*>(p);
  }
  void dfree(void* p) {
    if (!p) return;
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) allocs.erase(it);
    (void)hipFree(p);
  }
  Bank& B() { return bank[cur_bank]; }
  int* flag_in() { return d_flag[kseq % 3]; }
  int* flag_out() { return d_flag[(kseq + 1) % 3]; }
  int* flag_clear() { return d_flag[(kseq + 2) % 3]; }
};

namespace {

int fail(sc_system* s, int status, const std::string& msg) {
  if (s) s->err = msg;
  return status;
}

// ---- timing ---------------------------------------------------------------
hipEvent_t take_event(sc_system* s) {
  if (!s->ev_free.empty()) {
    hipEvent_t e = s->ev_free.back();
    s->ev_free.pop_back();
    return e;
  }
  hipEvent_t e;
  HIPCK(hipEventCreate(&e));
  return e;
}

void resolve_events(sc_system* s) {
  for (auto& p : s->ev_pending) {
    float ms = 0;
    HIPCK(hipEventElapsedTime(&ms, p.a, p.b));
    s->kt_ms[p.kernel] += ms;
    s->kt_n[p.kernel] += 1;
    s->ev_free.push_back(p.a);
    s->ev_free.push_back(p.b);
  }
  s->ev_pending.clear();
}

// A stream-captured sequence of launches, replayed with one call.  The
// capture runs the launch code once (advancing the host-side buffer rotation
// as the replays will); nothing in a captured sequence may be a memset.
struct GraphExec {
  sc_system* s;
  hipGraphExec_t exec = nullptr;
  explicit GraphExec(sc_system* sys) : s(sys) {}
  template <class F>
  void capture(F&& launches) {
    HIPCK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
      launches();
    } catch (...) {
      (void)hipStreamEndCapture(s->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    HIPCK(hipStreamEndCapture(s->stream, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      exec = nullptr;
      throw HipError{e, "hipGraphInstantiate"};
    }
  }
  void launch() { HIPCK(hipGraphLaunch(exec, s->stream)); }
  ~GraphExec() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};

struct KernelTimer {
  sc_system* s;
  int kernel;
  hipEvent_t a{}, b{};
  KernelTimer(sc_system* sys, int k) : s(sys), kernel(k) {
    if (!s->timing) return;
    if (s->ev_pending.size() >= 4096) {
      HIPCK(hipStreamSynchronize(s->stream));
      resolve_events(s);
    }
    a = take_event(s);
    b = take_event(s);
    HIPCK(hipEventRecord(a, s->stream));
  }
  ~KernelTimer() noexcept(false) {
    if (!s->timing) return;
    HIPCK(hipEventRecord(b, s->stream));
    s->ev_pending.push_back({a, b, kernel});
  }
};

// ---- geometry -------------------------------------------------------------

void build_tiles(sc_system* s) {
  s->tiles.clear();
  s->tile_ranges.assign(s->npack, make_int2(0, 0));
  int coff = 0;
  for (int k = 0; k < s->npack; ++k) {
    PackInfo& P = s->packs[k];
    P.coff = coff;
    coff += P.rows * P.cols;
    int tw = P.cols >= 3 ? std::min(s->tile_cols, P.cols - 2) : P.cols;
    tw = std::max(1, std::min(tw, kMaxTileCols));
    P.tile_cols = tw;
    P.cap = s->cap;
    std::vector<int4> spatial;
    for (int r = 0; r < P.rows; ++r)
      for (int c0 = 0; c0 < P.cols; c0 += tw)
        spatial.push_back(make_int4(k, r, c0, std::min(tw, P.cols - c0)));
    P.toff = (int)s->tiles.size();
    P.ntiles = (int)spatial.size();
    s->tile_ranges[k] = make_int2(P.toff, P.ntiles);
    auto launch = xcd_order(spatial, P.toff);
    s->tiles.insert(s->tiles.end(), launch.begin(), launch.end());
  }
  s->ncells = coff;
  s->ntiles = (int)s->tiles.size();
}

void build_pblocks(sc_system* s) {
  s->pblocks.clear();
  s->list_ranges.assign(s->npack, make_int2(0, 0));
  for (int k = 0; k < s->npack; ++k) {
    const PackInfo& P = s->packs[k];
    std::vector<int2> spatial;
    // list_lanes threads share a particle
    const int per_block = s->list_block / s->list_lanes;
    for (int st = P.poff; st < P.poff + P.n; st += per_block) spatial.push_back(make_int2(k, st));
    s->list_ranges[k] = make_int2((int)s->pblocks.size(), (int)spatial.size());
    auto launch = xcd_order(spatial, (int)s->pblocks.size());
    s->pblocks.insert(s->pblocks.end(), launch.begin(), launch.end());
  }
  s->npblocks = (int)s->pblocks.size();
}

template <typename T>
void upload_vec(sc_system* s, T*& d, const std::vector<T>& h) {
  s->dfree(d);
  d = s->dalloc<T>(h.size());
  HIPCK(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s->stream));
}

// (Re)allocate everything whose size depends on the cell grid / work maps.
void alloc_grid(sc_system* s) {
  for (int i = 0; i < 3; ++i) {
    s->dfree(s->d_cnt[i]);
    s->dfree(s->d_slot[i]);
    s->d_cnt[i] = s->dalloc<int>(s->ncells);
    s->d_slot[i] = s->dalloc<int>((size_t)s->ncells * s->cap);
    HIPCK(hipMemsetAsync(s->d_cnt[i], 0, sizeof(int) * s->ncells, s->stream));
  }
  s->dfree(s->d_cstart);
  s->d_cstart = s->dalloc<int>(s->ncells);
  upload_vec(s, s->d_tiles, s->tiles);
  upload_vec(s, s->d_tile_ranges, s->tile_ranges);
  s->dfree(s->d_tile_part);
  s->d_tile_part = s->dalloc<double>((size_t)s->ntiles * kNSum);
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
}

// Update workgroups: each covers up to 256 particles of one packing.
std::vector<int2> update_blocks(const sc_system* s, const std::vector<int>& packs) {
  std::vector<int2> ub;
  for (int k : packs) {
    const PackInfo& P = s->packs[k];
    for (int st = P.poff; st < P.poff + P.n; st += 256) ub.push_back(make_int2(k, st));
  }
  return ub;
}

void alloc_pblocks(sc_system* s) {
  build_pblocks(s);
  upload_vec(s, s->d_pblocks, s->pblocks);
  upload_vec(s, s->d_list_ranges, s->list_ranges);
  s->dfree(s->d_list_part);
  s->d_list_part = s->dalloc<double>((size_t)s->npblocks * kNSum);
}

Bank alloc_bank(sc_system* s) {
  Bank b{};
  size_t n = s->ntot;
  for (double** p : {&b.x, &b.y, &b.phi, &b.R, &b.A, &b.I, &b.RC, &b.fx, &b.fy, &b.tau, &b.vx,
                     &b.vy, &b.vpsi, &b.nx, &b.ny})
    *p = s->dalloc<double>(n);
  b.id = s->dalloc<int>(n);
  for (double* p : {b.fx, b.fy, b.tau, b.vx, b.vy, b.vpsi})
    HIPCK(hipMemsetAsync(p, 0, sizeof(double) * n, s->stream));
  return b;
}

void upload_scal(sc_system* s) {
  HIPCK(hipMemcpyAsync(s->d_scal[s->cur_scal], s->scal.data(), sizeof(PackScal) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
}

void download_scal(sc_system* s) {
  HIPCK(hipMemcpyAsync(s->scal.data(), s->d_scal[s->cur_scal], sizeof(PackScal) * s->npack,
                       hipMemcpyDeviceToHost, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
}

int check_errors(sc_system* s) {
  int e = 0;
  HIPCK(hipMemcpyAsync(&e, s->d_err, sizeof(int), hipMemcpyDeviceToHost, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  HIPCK(hipGetLastError());
  if (s->timing) resolve_events(s);
  if (!e) return SC_OK;
  HIPCK(hipMemsetAsync(s->d_err, 0, sizeof(int), s->stream));
  if (e & kErrCellOverflow)
    return fail(s, SC_ERR_CELL_OVERFLOW,
                "a cell holds more particles than its bucket (" + std::to_string(s->cap) + ")");
  if (e & kErrList)
    return fail(s, SC_ERR_LIST_OVERFLOW, "a particle has more near candidates than its list holds");
  return fail(s, SC_ERR_ARG, "non-finite particle coordinate");
}

int grid_for(int n, int block) { return (n + block - 1) / block; }

// Every neighbour-list row is rebuilt by the next force pass.
void force_rebuild(sc_system* s) {
  HIPCK(hipMemsetAsync(s->flag_in(), 1, sizeof(int) * s->npack, s->stream));
}

void segvec(sc_system* s) {
  hipLaunchKernelGGL(k_segvec, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->ntot);
}

// Bin every particle of the current bank into bucket buffer `buf` (cleared).
void bin_all(sc_system* s, int buf, int wrap_angles) {
  HIPCK(hipMemsetAsync(s->d_cnt[buf], 0, sizeof(int) * s->ncells, s->stream));
  hipLaunchKernelGGL(k_bin_all, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_packs, s->d_scal[s->cur_scal], s->d_cnt[buf],
                     s->d_slot[buf], s->d_err, s->ntot, wrap_angles);
}

// Raise the bucket capacity when bucket buffer `buf` (just filled by a binning
// of every particle) holds a cell with more than half the capacity: the new
// capacity is the next power of two at or above twice the largest count (at
// most kCapMax).  The slot arrays of all three buffers are reallocated (their
// counts stay valid: the buffer in use is binned again, the others are
// cleared before their next fill).  Returns whether the capacity grew.
bool grow_capacity(sc_system* s, int buf) {
  std::vector<int> h(s->ncells);
  HIPCK(hipMemcpyAsync(h.data(), s->d_cnt[buf], sizeof(int) * s->ncells, hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  int most = 0;
  for (int v : h) most = std::max(most, v);
  if (2 * most <= s->cap || s->cap >= kCapMax) return false;
  int cap = s->cap;
  while (cap < 2 * most && cap < kCapMax) cap *= 2;
  s->cap = cap;
  for (auto& P : s->packs) P.cap = cap;
  for (int i = 0; i < 3; ++i) {
    s->dfree(s->d_slot[i]);
    s->d_slot[i] = s->dalloc<int>((size_t)s->ncells * cap);
  }
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
  HIPCK(hipMemsetAsync(s->d_err, 0, sizeof(int), s->stream));  // the overflow it cures
  return true;
}

// bin_all, then raise the capacity and bin again if the buckets were crowded.
void bin_all_sized(sc_system* s, int buf, int wrap_angles) {
  bin_all(s, buf, wrap_angles);
  if (grow_capacity(s, buf)) bin_all(s, buf, 0);
}

// Counting sort of the current bank by the cells of bucket buffer cur_buf:
// per-packing scan of the counts, per-cell ordering by id, coalesced gather.
void resort(sc_system* s) {
  KernelTimer t(s, SC_KERNEL_SORT);
  int buf = s->cur_buf;
  hipLaunchKernelGGL(k_scan_cells, dim3(s->npack), dim3(256), 0, s->stream, s->d_packs,
                     s->d_cnt[buf], s->d_cstart);
  hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(s->ncells, 256)), dim3(256), 0, s->stream,
                     s->d_cstart, s->d_cnt[buf], s->d_slot[buf], s->B().id, s->d_perm, s->ncells,
                     s->cap);
  Bank& src = s->bank[s->cur_bank];
  Bank& dst = s->bank[1 - s->cur_bank];
  hipLaunchKernelGGL(k_gather, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, src, dst,
                     s->d_perm, s->ntot, s->velocity_live ? 1 : 0);
  s->cur_bank = 1 - s->cur_bank;
  force_rebuild(s);  // list rows hold storage indices
}

void ensure_se_ring(sc_system* s, size_t nrec) {
  size_t need = nrec * s->npack * kNSe;
  if (need <= s->se_ring_cap) return;
  s->dfree(s->d_se_ring);
  s->d_se_ring = s->dalloc<double>(need);
  s->se_ring_cap = need;
}

// ---- one force pass (either path) -----------------------------------------

// What a force pass does besides forces.  Positions and cells it evaluates:
// the current bank (pred, forces) or the predicted state (corr).
struct PassSpec {
  int mode = kModeForces;
  int se_every = 0, se_always = 0, do_fire = 0, do_gd = 0, do_resid = 0;
  int resize = 0, shrink = 0, corrected_image = 0, pure = 0;
  int read_buf = 0, fill_buf = -1, clear_buf = -1;
  int kernel_id = SC_KERNEL_FORCE;
};

void launch_force(sc_system* s, const PassSpec& ps) {
  const bool predicted = ps.mode == kModeCorr;
  if (s->path == SC_PATH_TILE) {
    TileArgs a{};
    a.packs = s->d_packs;
    a.scal = s->d_scal[s->cur_scal];
    a.scal_next = s->d_scal[1 - s->cur_scal];
    a.tiles = s->d_tiles;
    a.cnt = s->d_cnt[ps.read_buf];
    a.slot = s->d_slot[ps.read_buf];
    a.cnt_fill = ps.fill_buf >= 0 ? s->d_cnt[ps.fill_buf] : nullptr;
    a.slot_fill = ps.fill_buf >= 0 ? s->d_slot[ps.fill_buf] : nullptr;
    a.cnt_clear = ps.clear_buf >= 0 ? s->d_cnt[ps.clear_buf] : nullptr;
    a.ncells_total = s->ncells;
    a.b = s->B();
    a.sx = predicted ? s->d_px : a.b.x;
    a.sy = predicted ? s->d_py : a.b.y;
    a.sphi = predicted ? s->d_pphi : a.b.phi;
    a.px = s->d_px;
    a.py = s->d_py;
    a.pphi = s->d_pphi;
    a.v0x = s->d_v0x;
    a.v0y = s->d_v0y;
    a.v0t = s->d_v0t;
    a.tile_part = s->d_tile_part;
    a.err = s->d_err;
    a.se_every = ps.se_every;
    a.se_always = ps.se_always;
    a.do_fire = ps.do_fire;
    a.do_gd = ps.do_gd;
    a.do_resid = ps.do_resid;
    a.write_forces = 1;
    a.resize = ps.resize;
    a.shrink = ps.shrink;
    a.corrected_image = ps.corrected_image;
    a.two_point = s->contact_model == SC_CONTACT_TWO_POINT;
    a.npack = s->npack;
    a.list_groups = s->block / kLanes;
    a.stage_cap = 3 * (s->tile_cols + 2) * s->cap;
    size_t lds = tile_lds_bytes(a.stage_cap, a.list_groups, s->block);
    if (lds > (size_t)s->lds_max) throw StatusError{SC_ERR_ARG, "tile too wide for LDS"};
    using TFn = void (*)(TileArgs);
    static const TFn ttable[2][3] = {
        {&k_tile<kModeForces, false>, &k_tile<kModePred, false>, &k_tile<kModeCorr, false>},
        {&k_tile<kModeForces, true>, &k_tile<kModePred, true>, &k_tile<kModeCorr, true>}};
    TFn fn = ttable[a.two_point ? 1 : 0][ps.mode];
    if (lds > 64 * 1024)
      HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(fn),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    KernelTimer t(s, ps.kernel_id);
    hipLaunchKernelGGL(fn, dim3(s->ntiles), dim3(s->block), lds, s->stream, a);
  } else {
    ListArgs a{};
    a.packs = s->d_packs;
    a.scal = s->d_scal[s->cur_scal];
    a.scal_next = s->d_scal[1 - s->cur_scal];
    a.pblocks = s->d_pblocks;
    a.b = s->B();
    a.ex = predicted ? s->d_px : a.b.x;
    a.ey = predicted ? s->d_py : a.b.y;
    a.enx = predicted ? s->d_pnx : a.b.nx;
    a.eny = predicted ? s->d_pny : a.b.ny;
    a.px = s->d_px;
    a.py = s->d_py;
    a.pphi = s->d_pphi;
    a.pnx = s->d_pnx;
    a.pny = s->d_pny;
    a.v0x = s->d_v0x;
    a.v0y = s->d_v0y;
    a.v0t = s->d_v0t;
    a.cnt = s->d_cnt[ps.read_buf];
    a.slot = s->d_slot[ps.read_buf];
    a.cnt_fill = ps.fill_buf >= 0 ? s->d_cnt[ps.fill_buf] : nullptr;
    a.slot_fill = ps.fill_buf >= 0 ? s->d_slot[ps.fill_buf] : nullptr;
    a.cnt_clear = ps.clear_buf >= 0 ? s->d_cnt[ps.clear_buf] : nullptr;
    a.ncells_total = s->ncells;
    a.nbr = s->d_nbr;
    a.nbr_n = s->d_nbr_n;
    a.xb = s->d_xb;
    a.yb = s->d_yb;
    a.list_center = s->d_list_center;
    a.flag_in = s->flag_in();
    a.flag_out = s->flag_out();
    a.flag_clear = s->flag_clear();
    a.gamma_window = s->gamma_window;
    a.part = s->d_list_part;
    a.err = s->d_err;
    a.ntot = s->ntot;
    a.npack = s->npack;
    a.se_every = ps.se_every;
    a.se_always = ps.se_always;
    a.do_fire = ps.do_fire;
    a.do_gd = ps.do_gd;
    a.do_resid = ps.do_resid;
    a.resize = ps.resize;
    a.shrink = ps.shrink;
    a.corrected_image = ps.corrected_image;
    a.pure = ps.pure;
    a.two_point = s->contact_model == SC_CONTACT_TWO_POINT;
    using Fn = void (*)(ListArgs);
#define SC_LIST_ROW(M, T, X, E, B)                                                   \
  {&k_list<M, 1, T, X, E, B>, &k_list<M, 2, T, X, E, B>, &k_list<M, 4, T, X, E, B>, \
   &k_list<M, 8, T, X, E, B>}
#define SC_LIST_MODES(T, X)                                                            \
  {SC_LIST_ROW(kModeForces, T, X, true, false), SC_LIST_ROW(kModePred, T, X, false, false), \
   SC_LIST_ROW(kModeCorr, T, X, false, false), SC_LIST_ROW(kModeForces, T, X, false, false), \
   SC_LIST_ROW(kModePred, T, X, false, true), SC_LIST_ROW(kModeCorr, T, X, false, true)}
    // [pair precision: FP64 only / mixed][contact model][mode][lanes].  Modes
    // 0-2: force pass, sheared predictor, sheared corrector; 3: force pass
    // that records no stress / energy (minimiser passes); 4, 5: predictor and
    // corrector of steps that change the box (resize, pure shear)
    static const Fn table[2][2][6][4] = {
        {SC_LIST_MODES(false, false), SC_LIST_MODES(true, false)},
        {SC_LIST_MODES(false, true), SC_LIST_MODES(true, true)}};
#undef SC_LIST_MODES
#undef SC_LIST_ROW
    const int li = s->list_lanes == 1 ? 0 : (s->list_lanes == 2 ? 1 : (s->list_lanes == 4 ? 2 : 3));
    const bool se = ps.se_always || ps.se_every > 0;
    // Mixed precision runs the dynamics and the minimiser on FP32 pair forces;
    // a pass that records stress and energy evaluates its pairs in FP64 (the
    // FP64-only kernel), so reported stress and energy are FP64 evaluations of
    // the state the mixed run reached.
    const int xi = (s->prec == SC_PREC_MIXED && !(ps.mode == kModeForces && se)) ? 1 : 0;
    const bool box = ps.resize || ps.pure;
    const int mi = ps.mode == kModeForces ? (se ? 0 : 3) : (box ? ps.mode + 3 : ps.mode);
    Fn fn = table[xi][a.two_point ? 1 : 0][mi][li];
    KernelTimer t(s, ps.kernel_id);
    a.block_idx = s->compact ? s->d_act_blocks : nullptr;
    hipLaunchKernelGGL(fn, dim3(s->compact ? s->n_act_blocks : s->npblocks), dim3(s->list_block),
                       sizeof(int) * kHitCap * s->list_block, s->stream, a);
    s->kseq++;
  }
}

// Per-packing reduction of the last force pass's partial sums.  It reads
// scal[cur] and, for FIRE / gradient descent, writes the decisions to scal[next].
void launch_reduce(sc_system* s, const PassSpec& ps, const FireParams* fp) {
  ReduceArgs r{};
  r.packs = s->d_packs;
  r.scal = s->d_scal[s->cur_scal];
  r.scal_next = s->d_scal[1 - s->cur_scal];
  r.part = s->path == SC_PATH_TILE ? s->d_tile_part : s->d_list_part;
  r.ranges = s->path == SC_PATH_TILE ? s->d_tile_ranges : s->d_list_ranges;
  r.pack_sum = s->d_pack_sum;
  r.se_ring = s->d_se_ring;
  if (fp) r.fire = *fp;
  r.se_every = ps.se_every;
  r.se_always = ps.se_always;
  r.do_fire = ps.do_fire;
  r.do_gd = ps.do_gd;
  r.npack = s->npack;
  const bool compact = s->compact && s->path == SC_PATH_LIST;
  r.pack_idx = compact ? s->d_act_packs : nullptr;
  hipLaunchKernelGGL(k_reduce, dim3(compact ? s->n_act_packs : s->npack), dim3(256), 0, s->stream,
                     r);
}

// Plain force pass at the current state (calc_forces / calc_se), reduced.
void force_pass(sc_system* s, bool se, int se_every) {
  PassSpec ps;
  ps.mode = kModeForces;
  ps.se_always = se ? 1 : 0;
  ps.se_every = se_every;
  ps.do_resid = 1;
  ps.read_buf = s->cur_buf;
  launch_force(s, ps);
  launch_reduce(s, ps, nullptr);
}

// Zero the two bucket buffers that do not hold the current cells, so the
// first fill of a stepping loop starts from empty buckets.
void clear_other_buffers(sc_system* s) {
  for (int i = 1; i <= 2; ++i)
    HIPCK(hipMemsetAsync(s->d_cnt[(s->cur_buf + i) % 3], 0, sizeof(int) * s->ncells, s->stream));
}

void set_prec_all(sc_system* s, int prec) {
  for (auto& S : s->scal) S.prec = prec;
}

// One Heun sub-step (shear or resize) on every packing.  Bucket buffers
// rotate: c0 holds the cells of x, the predictor bins x~ into c1 and zeroes
// c2, the corrector reads c1, bins x' into c2 and zeroes c0.
void heun_substep(sc_system* s, bool resize, bool shrink, bool corrected, int se_every, int st,
                  bool pure = false) {
  const int c0 = s->cur_buf, c1 = (c0 + 1) % 3, c2 = (c0 + 2) % 3;
  if (se_every > 0 && st > 0 && st % se_every == 0) {
    // stress / energy record of the state the step starts from (the reference
    // evaluates it with calc_se after the previous step, strain_dynamics.cpp:289-291)
    PassSpec r;
    r.mode = kModeForces;
    r.se_every = se_every;
    r.read_buf = c0;
    launch_force(s, r);
    launch_reduce(s, r, nullptr);
  }
  PassSpec p;
  p.mode = kModePred;
  p.resize = resize;
  p.shrink = shrink;
  p.corrected_image = corrected;
  p.pure = pure;
  p.read_buf = c0;
  p.fill_buf = c1;
  p.clear_buf = c2;
  p.kernel_id = SC_KERNEL_PREDICT;
  launch_force(s, p);
  PassSpec c = p;
  c.mode = kModeCorr;
  c.read_buf = c1;
  c.fill_buf = c2;
  c.clear_buf = c0;
  c.kernel_id = SC_KERNEL_CORRECT;
  launch_force(s, c);
  s->cur_buf = c2;
  s->cur_scal = 1 - s->cur_scal;
}

// Restrict the minimiser's launches to the packings that have not finished
// (list path): their force workgroups, their reductions and their update
// workgroups.  Finished packings keep their last state; their cells are
// filled again before the final evaluation.
void set_compact(sc_system* s, const std::vector<int>& active) {
  std::vector<int> blocks, packs;
  for (int k : active) {
    const int2 r = s->list_ranges[k];
    for (int b = r.x; b < r.x + r.y; ++b) blocks.push_back(b);
    packs.push_back(k);
  }
  const std::vector<int2> ublocks = update_blocks(s, active);
  upload_vec(s, s->d_act_blocks, blocks);
  upload_vec(s, s->d_act_packs, packs);
  upload_vec(s, s->d_act_ublocks, ublocks);
  HIPCK(hipStreamSynchronize(s->stream));  // the host vectors go out of scope
  s->n_act_blocks = (int)blocks.size();
  s->n_act_packs = (int)packs.size();
  s->n_act_ublocks = (int)ublocks.size();
  s->compact = true;
}

// Every update workgroup reducing its packing's partials itself reads
// (update workgroups) x (force workgroups) x kNSum doubles per packing, which
// grows with N^2.  Up to this many bytes in all (and force workgroups per
// packing) the update kernel reduces; beyond it k_reduce reduces each packing
// once, in a launch of its own.
constexpr double kFuseMaxRereadBytes = 32.0 * 1024 * 1024;
// Lanes of the chip (256 CUs x 4 SIMDs x 64): the minimiser's tail widens the
// lanes per particle until its launch has about this many.
constexpr long kTailLanes = 65536;
constexpr int kFuseMaxBlocks = 4096;

// Gradient-descent (mode 0) / FIRE (mode 1) step after a force pass: the
// per-packing reduction and decision (fused into the update when every
// packing's partials are few, else a k_reduce launch first), then the update.
void launch_update(sc_system* s, const PassSpec& ps, const FireParams& fp, int mode) {
  const int b1 = (s->cur_buf + 1) % 3, b2 = (s->cur_buf + 2) % 3;
  const bool list = s->path == SC_PATH_LIST;
  const bool compact = s->compact && list;
  const std::vector<int2>& ranges = list ? s->list_ranges : s->tile_ranges;
  int most = 0;
  double reread = 0.0;
  for (int k = 0; k < s->npack; ++k) {
    most = std::max(most, ranges[k].y);
    reread += (double)((s->packs[k].n + 255) / 256) * ranges[k].y * kNSum * sizeof(double);
  }
  const bool fused = most <= kFuseMaxBlocks && reread <= kFuseMaxRereadBytes;
  if (!fused) launch_reduce(s, ps, &fp);
  UpdateArgs u{};
  u.b = s->B();
  u.pack_of = s->d_pack_of;
  u.packs = s->d_packs;
  ReduceArgs& r = u.red;
  r.packs = s->d_packs;
  r.scal = s->d_scal[s->cur_scal];
  r.scal_next = s->d_scal[1 - s->cur_scal];
  r.part = list ? s->d_list_part : s->d_tile_part;
  r.ranges = list ? s->d_list_ranges : s->d_tile_ranges;
  r.pack_sum = s->d_pack_sum;
  r.se_ring = s->d_se_ring;
  r.fire = fp;
  r.se_every = ps.se_every;
  r.se_always = ps.se_always;
  r.do_fire = ps.do_fire;
  r.do_gd = ps.do_gd;
  r.npack = s->npack;
  u.fused = fused ? 1 : 0;
  u.cnt = s->d_cnt[b1];
  u.slot = s->d_slot[b1];
  u.cnt_clear = s->d_cnt[b2];
  u.err = s->d_err;
  u.xb = s->d_xb;
  u.yb = s->d_yb;
  u.flag_out = list ? s->flag_out() : nullptr;
  u.flag_clear = list ? s->flag_clear() : nullptr;
  u.skin = 0;  // per packing: the cell padding (read in the kernel)
  u.max_move = mode == 1 ? fp.max_move : 0.0;
  u.ntot = s->ntot;
  u.npack = s->npack;
  u.mode = mode;
  u.ublocks = compact ? s->d_act_ublocks : s->d_ublocks;
  {
    KernelTimer t(s, SC_KERNEL_FIRE_UPDATE);
    hipLaunchKernelGGL(k_update, dim3(compact ? s->n_act_ublocks : (int)s->ublocks.size()),
                       dim3(256), 0, s->stream, u);
  }
  if (list) s->kseq++;
  s->cur_buf = b1;
  s->cur_scal = 1 - s->cur_scal;
}

// FIRE of the packings in `active` in one launch of k_fire_small (one
// workgroup per packing; every packing holds at most s->small_max <= kSmallMax
// particles), to convergence or for at most s->small_steps iterations.  Positions, velocities, forces and the
// per-packing scalars are written back in place; the rows of the list path no
// longer match the moved particles, so every row is rebuilt by the next pass.
void run_small_fire(sc_system* s, const std::vector<int>& active, const FireParams& fp) {
  if (!s->d_small_idx) s->d_small_idx = s->dalloc<int>(s->npack);
  HIPCK(hipMemcpyAsync(s->d_small_idx, active.data(), sizeof(int) * active.size(),
                       hipMemcpyHostToDevice, s->stream));
  SmallFireArgs a{};
  a.b = s->B();
  a.packs = s->d_packs;
  a.scal = s->d_scal[s->cur_scal];
  a.pack_idx = s->d_small_idx;
  a.err = s->d_err;
  a.fire = fp;
  // a launch stops after s->small_steps iterations (50000 by default, about
  // half a second), so no single kernel runs long; the minimiser's loop
  // launches again for packings still running, whose rows are rebuilt from
  // the state reached
  a.max_steps = s->small_steps;
  const bool tp = s->contact_model == SC_CONTACT_TWO_POINT;
  const bool mixed = s->prec == SC_PREC_MIXED;
  auto kern = tp ? (mixed ? &k_fire_small<true, true> : &k_fire_small<true, false>)
                 : (mixed ? &k_fire_small<false, true> : &k_fire_small<false, false>);
  {
    KernelTimer t(s, SC_KERNEL_FIRE_FORCE);
    hipLaunchKernelGGL(kern, dim3((unsigned)active.size()), dim3(kSmallThreads), 0, s->stream,
                       a);
  }
  force_rebuild(s);
}

int reconfigure_if_needed(sc_system* s, const std::vector<std::pair<double, double>>* Lhint);
bool regrid(sc_system* s);
// systems up to this many particles step through captured graphs
constexpr int kGraphMaxParticles = 65536;

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int sc_version(void) { return 10100; }

const char* sc_status_string(int status) {
  switch (status) {
    case SC_OK: return "ok";
    case SC_ERR_ARG: return "invalid argument";
    case SC_ERR_HIP: return "HIP runtime error";
    case SC_ERR_NO_DEVICE: return "no GPU device";
    case SC_ERR_CELL_OVERFLOW: return "cell bucket overflow";
    case SC_ERR_PLACEMENT: return "random placement failed";
    case SC_ERR_LIST_OVERFLOW: return "candidate list overflow";
    case SC_ERR_NOT_CONVERGED: return "minimiser did not converge";
    case SC_ERR_STATE: return "invalid state";
    case SC_ERR_HOST: return "host memory or internal error";
    default: return "unknown status";
  }
}

int sc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

const char* sc_last_error(const sc_system* s) { return s ? s->err.c_str() : ""; }

int sc_num_particles(const sc_system* s, int* n_total, int* npack) {
  if (!s) return SC_ERR_ARG;
  if (n_total) *n_total = s->ntot;
  if (npack) *npack = s->npack;
  return SC_OK;
}

int sc_system_create(sc_system** out, int device, int npack, const int* offsets, const double* L,
                     const double* x, const double* y, const double* phi, const double* A,
                     const double* R, const double* pad) {
  if (!out || npack <= 0 || !offsets || !L || !x || !y || !phi || !A || !R || !pad)
    return SC_ERR_ARG;
  *out = nullptr;
  if (offsets[0] != 0) return SC_ERR_ARG;
  for (int k = 0; k < npack; ++k)
    if (offsets[k + 1] <= offsets[k] || !(L[k] > 0) || !(pad[k] >= 0)) return SC_ERR_ARG;
  int ntot = offsets[npack];
  for (int i = 0; i < ntot; ++i)
    if (!(R[i] > 0) || !(A[i] >= 0) || !std::isfinite(x[i]) || !std::isfinite(y[i]) ||
        !std::isfinite(phi[i]))
      return SC_ERR_ARG;
  int ndev = sc_device_count();
  if (ndev <= 0) return SC_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return SC_ERR_ARG;

  sc_system* s = new (std::nothrow) sc_system();
  if (!s) return SC_ERR_HOST;
  try {
    s->device = device;
    HIPCK(hipSetDevice(device));
    HIPCK(hipDeviceGetAttribute(&s->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    HIPCK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    s->npack = npack;
    s->ntot = ntot;
    s->offsets.assign(offsets, offsets + npack + 1);
    s->packs.resize(npack);
    s->scal.assign(npack, PackScal{});
    s->Rmax.assign(npack, 0.0);
    s->Amax.assign(npack, 0.0);
    s->filled.assign(npack, 0.0);
    std::vector<double> hx(x, x + ntot), hy(y, y + ntot), hp(phi, phi + ntot), hR(R, R + ntot),
        hA(A, A + ntot), hI(ntot), hRC(ntot);
    std::vector<int> hid(ntot), hpack(ntot);
    for (int k = 0; k < npack; ++k) {
      PackInfo& P = s->packs[k];
      P.poff = offsets[k];
      P.n = offsets[k + 1] - offsets[k];
      P.dt = 0.05;  // SpherocylBox::set_defaults, src/cpp/spherocyl_box.cpp:24
      P.strain_rate = 0;
      P.resize_rate = 0;
      P.resize_scale = 1.0;
      P.inv_2n = 1.0 / (2.0 * P.n);
      double filled = 0;
      for (int i = P.poff; i < P.poff + P.n; ++i) {
        // set_particle, src/cpp/spherocyl_box.cpp:198-219
        double al = hA[i] / hR[i];
        double C = (3 * kRefPi + 24 * al + 6 * kRefPi * al * al + 8 * al * al * al);
        double area = hR[i] * (kRefPi * hR[i] + 4 * hA[i]);
        hI[i] = hR[i] * hR[i] * hR[i] * hR[i] * C / (6 * area);
        hRC[i] = (8 * al + 6 * kRefPi * al * al + 8 * al * al * al) / C;
        while (hx[i] > L[k]) hx[i] -= L[k];
        while (hx[i] < 0.0) hx[i] += L[k];
        while (hy[i] > L[k]) hy[i] -= L[k];
        while (hy[i] < 0.0) hy[i] += L[k];
        while (hp[i] > kRefPi) hp[i] -= 2 * kRefPi;
        while (hp[i] < -kRefPi) hp[i] += 2 * kRefPi;
        s->Amax[k] = std::max(s->Amax[k], hA[i]);
        s->Rmax[k] = std::max(s->Rmax[k], hR[i]);
        filled += (kRefPi * hR[i] + 4.0 * hA[i]) * hR[i];
        hid[i] = i - P.poff;
        hpack[i] = k;
      }
      s->filled[k] = filled;
      P.pad = pad[k];
      cell_grid(L[k], L[k], s->Rmax[k], s->Amax[k], P.pad, P.rows, P.cols);
      PackScal& S = s->scal[k];
      S.Lx = S.Ly = L[k];
      S.prec = kFP64;
    }
    build_tiles(s);

    s->d_packs = s->dalloc<PackInfo>(npack);
    s->d_scal[0] = s->dalloc<PackScal>(npack);
    s->d_scal[1] = s->dalloc<PackScal>(npack);
    s->d_pack_sum = s->dalloc<double>((size_t)npack * kNSum);
    HIPCK(hipMemsetAsync(s->d_pack_sum, 0, sizeof(double) * npack * kNSum, s->stream));
    s->d_err = s->dalloc<int>(1);
    HIPCK(hipMemsetAsync(s->d_err, 0, sizeof(int), s->stream));
    s->bank[0] = alloc_bank(s);
    s->bank[1] = alloc_bank(s);
    for (double** p : {&s->d_px, &s->d_py, &s->d_pphi, &s->d_pnx, &s->d_pny, &s->d_v0x, &s->d_v0y,
                       &s->d_v0t, &s->d_xb, &s->d_yb})
      *p = s->dalloc<double>(ntot);
    s->d_pack_of = s->dalloc<int>(ntot);
    s->d_perm = s->dalloc<int>(ntot);
    HIPCK(hipMemsetAsync(s->d_perm, 0, sizeof(int) * ntot, s->stream));
    s->d_io = s->dalloc<double>((size_t)3 * ntot);
    s->d_scale = s->dalloc<double>(npack);
    s->d_nbr = s->dalloc<int>((size_t)kNbrCap * ntot);
    s->d_nbr_n = s->dalloc<int>(ntot);
    HIPCK(hipMemsetAsync(s->d_nbr_n, 0, sizeof(int) * ntot, s->stream));
    s->d_list_center = s->dalloc<double>(npack);
    for (int i = 0; i < 3; ++i) {
      s->d_flag[i] = s->dalloc<int>(npack);
      HIPCK(hipMemsetAsync(s->d_flag[i], 0, sizeof(int) * npack, s->stream));
    }
    alloc_grid(s);
    alloc_pblocks(s);
    {
      std::vector<int> all(npack);
      for (int k = 0; k < npack; ++k) all[k] = k;
      s->ublocks = update_blocks(s, all);
      upload_vec(s, s->d_ublocks, s->ublocks);
    }
    upload_scal(s);

    Bank& b = s->bank[0];
    auto up = [&](double* d, const std::vector<double>& h) {
      HIPCK(hipMemcpyAsync(d, h.data(), sizeof(double) * ntot, hipMemcpyHostToDevice, s->stream));
    };
    up(b.x, hx);
    up(b.y, hy);
    up(b.phi, hp);
    up(b.R, hR);
    up(b.A, hA);
    up(b.I, hI);
    up(b.RC, hRC);
    HIPCK(hipMemcpyAsync(b.id, hid.data(), sizeof(int) * ntot, hipMemcpyHostToDevice, s->stream));
    HIPCK(hipMemcpyAsync(s->d_pack_of, hpack.data(), sizeof(int) * ntot, hipMemcpyHostToDevice,
                         s->stream));
    s->cur_bank = 0;
    s->cur_buf = 0;
    segvec(s);
    bin_all_sized(s, 0, 0);
    resort(s);
    HIPCK(hipStreamSynchronize(s->stream));
    int st = check_errors(s);
    if (st != SC_OK) {
      sc_system_destroy(s);
      return st;
    }
  } catch (const HipError& e) {
    std::fprintf(stderr, "spherocyls: %s failed: %s\n", e.what, hipGetErrorString(e.e));
    sc_system_destroy(s);
    return SC_ERR_HIP;
  } catch (...) {
    std::string m;
    const int st = current_exception_status(&m);
    std::fprintf(stderr, "spherocyls: sc_system_create: %s\n", m.c_str());
    sc_system_destroy(s);
    return st;
  }
  *out = s;
  return SC_OK;
}

int sc_system_destroy(sc_system* s) {
  if (!s) return SC_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (auto& p : s->ev_pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : s->ev_free) (void)hipEventDestroy(e);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return SC_OK;
}

#define API_BEGIN            \
  if (!s) return SC_ERR_ARG; \
  try {                      \
    HIPCK(hipSetDevice(s->device));
#define API_END                                                                       \
  }                                                                                   \
  catch (const HipError& e) {                                                         \
    return fail(s, SC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e)); \
  }                                                                                   \
  catch (...) {                                                                       \
    std::string m_;                                                                   \
    const int st_ = current_exception_status(&m_);                                    \
    return fail(s, st_, m_);                                                          \
  }

int sc_set_dynamics(sc_system* s, int k, double strain_rate, double resize_rate, double step) {
  if (!s || k < -1 || k >= s->npack || !(step > 0) || !std::isfinite(strain_rate) ||
      !(resize_rate >= 0 && resize_rate < 1))
    return SC_ERR_ARG;
  API_BEGIN
  for (int i = 0; i < s->npack; ++i) {
    if (k >= 0 && i != k) continue;
    s->packs[i].strain_rate = strain_rate;
    s->packs[i].resize_rate = resize_rate;
    s->packs[i].dt = step;
    s->packs[i].resize_scale = std::pow(1.0 - resize_rate, step);
  }
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_set_gamma(sc_system* s, int k, double gamma, double gamma_total) {
  if (!s || k < -1 || k >= s->npack || !std::isfinite(gamma)) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  for (int i = 0; i < s->npack; ++i) {
    if (k >= 0 && i != k) continue;
    s->scal[i].gamma = gamma;
    s->scal[i].gamma_total = gamma_total;
  }
  upload_scal(s);
  force_rebuild(s);  // the lists' strain window moved
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_set_precision(sc_system* s, int mode) {
  if (!s || (mode != SC_PREC_FP64 && mode != SC_PREC_MIXED)) return SC_ERR_ARG;
  API_BEGIN
  s->prec = mode;
  download_scal(s);
  set_prec_all(s, mode == SC_PREC_FP64 ? kFP64 : kFP32);
  upload_scal(s);
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_set_contact_model(sc_system* s, int model) {
  if (!s || (model != SC_CONTACT_CLOSEST && model != SC_CONTACT_TWO_POINT)) return SC_ERR_ARG;
  s->contact_model = model;
  return SC_OK;
}

int sc_get_contact_model(const sc_system* s, int* model) {
  if (!s || !model) return SC_ERR_ARG;
  *model = s->contact_model;
  return SC_OK;
}

int sc_set_graphs(sc_system* s, int enable) {
  if (!s) return SC_ERR_ARG;
  s->use_graphs = enable != 0;
  return SC_OK;
}

int sc_set_small_fire(sc_system* s, int max_particles, int steps_per_launch) {
  if (!s || max_particles < 0 || max_particles > kSmallMax || steps_per_launch < 0)
    return SC_ERR_ARG;
  s->small_max = max_particles;
  s->small_steps = steps_per_launch > 0 ? steps_per_launch : 50000;
  return SC_OK;
}

int sc_set_path(sc_system* s, int path) {
  if (!s || (path != SC_PATH_TILE && path != SC_PATH_LIST)) return SC_ERR_ARG;
  API_BEGIN
  if (path != s->path) {
    s->path = path;
    segvec(s);  // the tile path does not keep (nx, ny) current
    force_rebuild(s);
  }
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_get_path(const sc_system* s, int* path) {
  if (!s || !path) return SC_ERR_ARG;
  *path = s->path;
  return SC_OK;
}

int sc_set_tiling(sc_system* s, int tile_cols, int block_threads) {
  if (!s || tile_cols < 1 || tile_cols > kMaxTileCols) return SC_ERR_ARG;
  if (block_threads != 64 && block_threads != 128 && block_threads != 256) return SC_ERR_ARG;
  if (tile_lds_bytes(3 * (tile_cols + 2) * s->cap, block_threads / kLanes, block_threads) >
      (size_t)s->lds_max)
    return fail(s, SC_ERR_ARG, "tile too wide for the workgroup's LDS");
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  s->tile_cols = tile_cols;
  s->block = block_threads;
  build_tiles(s);
  upload_vec(s, s->d_tiles, s->tiles);
  upload_vec(s, s->d_tile_ranges, s->tile_ranges);
  s->dfree(s->d_tile_part);
  s->d_tile_part = s->dalloc<double>((size_t)s->ntiles * kNSum);
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_get_tiling(const sc_system* s, int* tile_cols, int* block_threads) {
  if (!s) return SC_ERR_ARG;
  if (tile_cols) *tile_cols = s->path == SC_PATH_TILE ? s->tile_cols : s->list_lanes;
  if (block_threads) *block_threads = s->path == SC_PATH_TILE ? s->block : s->list_block;
  return SC_OK;
}

int sc_set_list_block(sc_system* s, int block_threads, int lanes) {
  if (!s || (block_threads != 64 && block_threads != 128 && block_threads != 256))
    return SC_ERR_ARG;
  if (lanes != 1 && lanes != 2 && lanes != 4 && lanes != 8) return SC_ERR_ARG;
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  s->list_block = block_threads;
  s->list_lanes = lanes;
  alloc_pblocks(s);
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_set_positions(sc_system* s, const double* x, const double* y, const double* phi) {
  if (!s || !x || !y || !phi) return SC_ERR_ARG;
  for (int i = 0; i < s->ntot; ++i)
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !std::isfinite(phi[i])) return SC_ERR_ARG;
  API_BEGIN
  size_t nb = sizeof(double) * s->ntot;
  HIPCK(hipMemcpyAsync(s->d_io, x, nb, hipMemcpyHostToDevice, s->stream));
  HIPCK(hipMemcpyAsync(s->d_io + s->ntot, y, nb, hipMemcpyHostToDevice, s->stream));
  HIPCK(hipMemcpyAsync(s->d_io + 2 * s->ntot, phi, nb, hipMemcpyHostToDevice, s->stream));
  hipLaunchKernelGGL(k_from_caller, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_packs, s->d_io, s->d_io + s->ntot, s->d_io + 2 * s->ntot,
                     s->ntot);
  s->cur_buf = 0;
  bin_all_sized(s, 0, 1);
  segvec(s);
  resort(s);
  return check_errors(s);
  API_END
}

static int to_caller(sc_system* s, int what, double* o0, double* o1, double* o2) {
  hipLaunchKernelGGL(k_to_caller, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_packs, what, s->d_io, s->d_io + s->ntot,
                     s->d_io + 2 * s->ntot, s->ntot);
  size_t nb = sizeof(double) * s->ntot;
  if (o0) HIPCK(hipMemcpyAsync(o0, s->d_io, nb, hipMemcpyDeviceToHost, s->stream));
  if (o1) HIPCK(hipMemcpyAsync(o1, s->d_io + s->ntot, nb, hipMemcpyDeviceToHost, s->stream));
  if (o2) HIPCK(hipMemcpyAsync(o2, s->d_io + 2 * s->ntot, nb, hipMemcpyDeviceToHost, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
}

int sc_get_positions(sc_system* s, double* x, double* y, double* phi) {
  API_BEGIN
  return to_caller(s, 0, x, y, phi);
  API_END
}

int sc_get_forces(sc_system* s, double* fx, double* fy, double* tau) {
  API_BEGIN
  return to_caller(s, 1, fx, fy, tau);
  API_END
}

int sc_get_box(sc_system* s, int k, double* out) {
  if (!s || !out || k < 0 || k >= s->npack) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  double ps[kNSum];
  HIPCK(hipMemcpy(ps, s->d_pack_sum + (size_t)k * kNSum, sizeof(ps), hipMemcpyDeviceToHost));
  const PackScal& S = s->scal[k];
  const PackInfo& P = s->packs[k];
  out[SC_BOX_LX] = S.Lx;
  out[SC_BOX_LY] = S.Ly;
  out[SC_BOX_GAMMA] = S.gamma;
  out[SC_BOX_GAMMA_TOTAL] = S.gamma_total;
  out[SC_BOX_E] = ps[kSeE];
  out[SC_BOX_PXX] = ps[kSePxx];
  out[SC_BOX_PYY] = ps[kSePyy];
  out[SC_BOX_PXY] = ps[kSePxy];
  out[SC_BOX_CPXX] = ps[kSeCPxx];
  out[SC_BOX_CPYY] = ps[kSeCPyy];
  out[SC_BOX_CPXY] = ps[kSeCPxy];
  out[SC_BOX_CPYX] = ps[kSeCPyx];
  out[SC_BOX_NCONTACTS] = ps[kSeNc];
  out[SC_BOX_FRMS] = std::sqrt(ps[kSumR2] / (3.0 * P.n));
  out[SC_BOX_PACKING] = s->filled[k] / (S.Lx * S.Ly);  // calculate_packing_fraction
  out[SC_BOX_ROWS] = P.rows;
  out[SC_BOX_COLS] = P.cols;
  out[SC_BOX_PAD] = P.pad;
  out[SC_BOX_ITERS] = S.iters;
  out[SC_BOX_FIRE_DT] = S.fire_dt;
  out[SC_BOX_FIRE_ALPHA] = S.fire_alpha;
  out[SC_BOX_PAIR_FP32] = S.prec == kFP32 ? 1.0 : 0.0;
  return SC_OK;
  API_END
}

int sc_calc_forces(sc_system* s) {
  API_BEGIN
  force_pass(s, false, 0);
  return check_errors(s);
  API_END
}

int sc_calc_se(sc_system* s) {
  API_BEGIN
  force_pass(s, true, 0);
  return check_errors(s);
  API_END
}

static void reset_step_counters(sc_system* s) {
  download_scal(s);
  for (auto& S : s->scal) {
    S.step = 0;
    S.converged = 0;
  }
  upload_scal(s);
}

static void copy_se_out(sc_system* s, int nrec, double* se_out) {
  if (se_out && nrec > 0)
    HIPCK(hipMemcpyAsync(se_out, s->d_se_ring, sizeof(double) * nrec * s->npack * kNSe,
                         hipMemcpyDeviceToHost, s->stream));
}

int sc_strain_steps(sc_system* s, int nsteps, int se_every, double* se_out) {
  if (!s || nsteps < 0 || se_every < 0) return SC_ERR_ARG;
  API_BEGIN
  int nrec = se_every > 0 ? nsteps / se_every : 0;
  if (nrec > 0) ensure_se_ring(s, nrec);
  reset_step_counters(s);
  clear_other_buffers(s);
  // Small systems are launch-bound: their steps run as a captured HIP graph of
  // 48 steps, replayed.  The bucket, flag and scalar buffers rotate with
  // periods 3, 3 and 2 steps, so 48 steps end where they began and one capture
  // serves every chunk.  Graphs skip the locality re-sort (irrelevant at this
  // size), stress records (their passes fall on particular steps) and event
  // timing; the steps left over run eagerly.
  const int resort_every = 1000;  // locality only: cells are exact every pass
  int st = 0;
  if (s->use_graphs && !s->timing && se_every == 0 && s->ntot <= kGraphMaxParticles) {
    const int chunk = 48;
    GraphExec g(s);
    for (; st + chunk <= nsteps; st += chunk) {
      if (!g.exec)
        g.capture([&]() {
          for (int m = 0; m < chunk; ++m) heun_substep(s, false, false, false, 0, m);
        });
      g.launch();
    }
  }
  for (; st < nsteps; ++st) {
    heun_substep(s, false, false, false, se_every, st);
    if ((st + 1) % resort_every == 0 && st + 1 < nsteps) {
      if (grow_capacity(s, s->cur_buf)) bin_all(s, s->cur_buf, 0);
      resort(s);
    }
  }
  force_pass(s, false, se_every);
  copy_se_out(s, nrec, se_out);
  int e = check_errors(s);
  download_scal(s);
  return e;
  API_END
}

int sc_resize_steps(sc_system* s, int nsteps, int shrink, int corrected_image, int se_every,
                    double* se_out) {
  if (!s || nsteps < 0 || se_every < 0) return SC_ERR_ARG;
  API_BEGIN
  int nrec = se_every > 0 ? nsteps / se_every : 0;
  if (nrec > 0) ensure_se_ring(s, nrec);
  reset_step_counters(s);
  clear_other_buffers(s);
  // the host mirrors L (same arithmetic as the device) to know when the cell
  // grid has to change without a round trip per step
  std::vector<std::pair<double, double>> Lm(s->npack);
  for (int k = 0; k < s->npack; ++k) Lm[k] = {s->scal[k].Lx, s->scal[k].Ly};
  for (int st = 0; st < nsteps; ++st) {
    // each packing resizes by its own (1 - rate)^dt
    heun_substep(s, true, shrink != 0, corrected_image != 0, se_every, st);
    for (int k = 0; k < s->npack; ++k) {
      const double scale = s->packs[k].resize_scale;
      auto& v = Lm[k];
      v.first = shrink ? v.first * scale : v.first / scale;
      v.second = shrink ? v.second * scale : v.second / scale;
    }
    int st2 = reconfigure_if_needed(s, &Lm);
    if (st2 != SC_OK) return st2;
  }
  force_pass(s, false, se_every);
  copy_se_out(s, nrec, se_out);
  int e = check_errors(s);
  download_scal(s);
  return e;
  API_END
}

int sc_pure_strain_steps(sc_system* s, int nsteps, int se_every, double* se_out) {
  if (!s || nsteps < 0 || se_every < 0) return SC_ERR_ARG;
  if (s->path != SC_PATH_LIST)
    return fail(s, SC_ERR_STATE, "pure shear runs on the list path (sc_set_path)");
  API_BEGIN
  int nrec = se_every > 0 ? nsteps / se_every : 0;
  if (nrec > 0) ensure_se_ring(s, nrec);
  reset_step_counters(s);
  clear_other_buffers(s);
  // host mirror of the box (the device's arithmetic) for the cell-grid checks
  std::vector<std::pair<double, double>> Lm(s->npack);
  for (int k = 0; k < s->npack; ++k) Lm[k] = {s->scal[k].Lx, s->scal[k].Ly};
  for (int st = 0; st < nsteps; ++st) {
    heun_substep(s, false, false, false, se_every, st, true);
    for (int k = 0; k < s->npack; ++k) {
      const double h = 0.5 * s->packs[k].dt * s->packs[k].strain_rate;
      Lm[k].first = Lm[k].first + h * Lm[k].first;
      Lm[k].second = Lm[k].second - h * Lm[k].second;
    }
    int st2 = reconfigure_if_needed(s, &Lm);
    if (st2 != SC_OK) return st2;
  }
  force_pass(s, false, se_every);
  copy_se_out(s, nrec, se_out);
  int e = check_errors(s);
  download_scal(s);
  return e;
  API_END
}

int sc_flip_shear(sc_system* s, int k) {
  if (!s || k < -1 || k >= s->npack) return SC_ERR_ARG;
  API_BEGIN
  std::vector<double> x(s->ntot), y(s->ntot), phi(s->ntot);
  int st = sc_get_positions(s, x.data(), y.data(), phi.data());
  if (st != SC_OK) return st;
  download_scal(s);
  std::vector<double> gam(s->npack);
  for (int i = 0; i < s->npack; ++i) {
    gam[i] = s->scal[i].gamma;
    if (k >= 0 && i != k) continue;
    const PackInfo& P = s->packs[i];
    const double Lx = s->scal[i].Lx;
    for (int g = P.poff; g < P.poff + P.n; ++g) {
      x[g] = Lx - x[g];
      phi[g] = -phi[g];
    }
    gam[i] = -gam[i];
  }
  st = sc_set_positions(s, x.data(), y.data(), phi.data());
  for (int i = 0; st == SC_OK && i < s->npack; ++i)
    if (k < 0 || i == k) st = sc_set_gamma(s, i, gam[i], s->scal[i].gamma_total);
  return st;
  API_END
}

int sc_rotate_by_gamma(sc_system* s, int k) {
  if (!s || k < -1 || k >= s->npack) return SC_ERR_ARG;
  API_BEGIN
  std::vector<double> x(s->ntot), y(s->ntot), phi(s->ntot);
  int st = sc_get_positions(s, x.data(), y.data(), phi.data());
  if (st != SC_OK) return st;
  download_scal(s);
  for (int i = 0; i < s->npack; ++i) {
    if (k >= 0 && i != k) continue;
    // Spherocyl_Box::rotate_by_gamma / gamma_rotate (src/cuda/find_neighbors.cu:
    // 580-645): rotate the lab frame so that the box vector (gamma Ly, Ly)
    // becomes the new x axis; the box becomes Ly sqrt(1+gamma^2) by
    // Lx / sqrt(1+gamma^2) at strain -gamma.
    PackScal& S = s->scal[i];
    const double g = S.gamma, root = std::sqrt(1.0 + g * g);
    const double sgn = std::signbit(g) ? -1.0 : 1.0;
    const double Lx = S.Lx, Ly = S.Ly;
    const double nLx = Ly * root, nLy = Lx / root;
    const double dphi = g != 0.0 ? std::atan(1.0 / g) : 0.5 * kRefPi;
    const PackInfo& P = s->packs[i];
    for (int j = P.poff; j < P.poff + P.n; ++j) {
      const double nx = sgn * y[j] * root, ny = -sgn * x[j] / root;
      double np = phi[j] - dphi;
      if (np > kRefPi) np -= 2 * kRefPi;
      else if (np < -kRefPi) np += 2 * kRefPi;
      x[j] = nx < 0 ? nx + nLx : (nx > nLx ? nx - nLx : nx);
      y[j] = ny < 0 ? ny + nLy : (ny > nLy ? ny - nLy : ny);
      phi[j] = np;
    }
    S.Lx = nLx;
    S.Ly = nLy;
    S.gamma = -g;
  }
  upload_scal(s);
  regrid(s);  // the cell grid of the new box; set_positions bins and re-sorts
  return sc_set_positions(s, x.data(), y.data(), phi.data());
  API_END
}

int sc_gd_steps(sc_system* s, int nmax, double emin, int* steps_done) {
  if (!s || nmax < 0) return SC_ERR_ARG;
  API_BEGIN
  reset_step_counters(s);
  clear_other_buffers(s);
  FireParams fp{};
  fp.emin = emin;
  fp.max_iter = nmax;
  // nmax updates need nmax + 1 checked force passes; chunks never launch more
  const int chunk = std::min(64, nmax + 1);
  while (true) {
    for (int m = 0; m < chunk; ++m) {
      PassSpec ps;
      ps.mode = kModeForces;
      ps.do_gd = 1;
      ps.se_always = 1;
      ps.read_buf = s->cur_buf;
      ps.kernel_id = SC_KERNEL_FIRE_FORCE;
      launch_force(s, ps);
      launch_update(s, ps, fp, 0);
    }
    int e = check_errors(s);
    if (e != SC_OK) return e;
    download_scal(s);
    bool all = true;
    for (auto& S : s->scal) all &= S.converged != 0;
    if (all) break;
  }
  int done = 0;
  for (auto& S : s->scal) done = std::max(done, S.iters);
  if (steps_done) *steps_done = done;
  // positions are final; leave forces, energy and stress of the final state
  // (the updates skip finished packings: fill every packing's cells first)
  s->cur_buf = 0;
  bin_all_sized(s, 0, 0);
  force_pass(s, true, 0);
  return check_errors(s);
  API_END
}

static int affine_resize(sc_system* s, const double* scale, bool y_only) {
  if (!s || !scale) return SC_ERR_ARG;
  for (int k = 0; k < s->npack; ++k)
    if (!(scale[k] > 0) || !std::isfinite(scale[k])) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  for (int k = 0; k < s->npack; ++k) {
    if (!y_only) s->scal[k].Lx *= scale[k];
    s->scal[k].Ly *= scale[k];
  }
  upload_scal(s);
  HIPCK(hipMemcpyAsync(s->d_scale, scale, sizeof(double) * s->npack, hipMemcpyHostToDevice,
                       s->stream));
  hipLaunchKernelGGL(k_affine, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_scale, s->ntot, y_only ? 1 : 0);
  int e = reconfigure_if_needed(s, nullptr);
  if (e != SC_OK) return e;
  s->cur_buf = 0;
  bin_all_sized(s, 0, 0);
  resort(s);
  return check_errors(s);
  API_END
}

int sc_affine_resize(sc_system* s, const double* scale) { return affine_resize(s, scale, false); }

int sc_affine_resize_y(sc_system* s, const double* scale) { return affine_resize(s, scale, true); }

int sc_fire_minimize(sc_system* s, const double* params, double switch_frms, int* iters,
                     double* frms) {
  if (!s || !params) return SC_ERR_ARG;
  API_BEGIN
  FireParams fp{};
  fp.dt0 = params[0];
  fp.dt_max = params[1];
  fp.n_min = params[2];
  fp.f_inc = params[3];
  fp.f_dec = params[4];
  fp.alpha0 = params[5];
  fp.f_alpha = params[6];
  fp.max_move = params[7];
  fp.ftol = params[8];
  fp.max_iter = params[9];
  fp.switch_frms = switch_frms;
  if (!(fp.dt0 > 0) || !(fp.dt_max >= fp.dt0) || !(fp.max_iter >= 0) || !(fp.max_move > 0))
    return fail(s, SC_ERR_ARG, "invalid FIRE parameters");
  for (auto& P : s->packs)
    if (!(fp.max_move < 0.5 * P.pad)) return fail(s, SC_ERR_ARG, "FIRE max_move must be < pad/2");
  download_scal(s);
  for (auto& S : s->scal) {
    S.step = 0;
    S.converged = 0;
    S.iters = 0;
    S.fire_dt = fp.dt0;
    S.fire_alpha = fp.alpha0;
    S.fire_npos = 0;
    S.prec = s->prec == SC_PREC_MIXED ? kFP32 : kFP64;
  }
  upload_scal(s);
  clear_other_buffers(s);
  for (int b = 0; b < 2; ++b)
    for (double* p : {s->bank[b].vx, s->bank[b].vy, s->bank[b].vpsi})
      HIPCK(hipMemsetAsync(p, 0, sizeof(double) * s->ntot, s->stream));
  s->velocity_live = true;
  // One FIRE iteration is two launches (force pass; reduction, decision and
  // update), or three for packings of many force workgroups, whose reduction
  // runs once in its own launch.  The rotating bucket, flag and scalar buffers repeat every 6
  // iterations, so a chunk of 30 is captured once as a HIP graph and replayed;
  // finished packings make every kernel return early.  Event timing needs
  // eager launches, and eager chunks re-sort the bank for locality.
  const int chunk = 30;
  const int resort_every = 240;
  const bool graph = s->use_graphs && !s->timing;
  GraphExec g(s);
  int it = 0;
  auto iterations = [&](int count) {
    for (int m = 0; m < count; ++m) {
      PassSpec ps;
      ps.mode = kModeForces;
      ps.do_fire = 1;
      ps.read_buf = s->cur_buf;
      ps.kernel_id = SC_KERNEL_FIRE_FORCE;
      launch_force(s, ps);
      launch_update(s, ps, fp, 1);
    }
  };
  // Once half of the packings still running have finished, later launches
  // cover only the rest (list path): the long tail of a few slow packings then
  // costs their own work, not a pass over every packing.  The full lists come
  // back however this function exits.
  struct CompactOff {
    sc_system* s;
    ~CompactOff() { s->compact = false; }
  } compact_off{s};
  int last_active = s->npack;
  // the system's own list shape, restored before the final evaluation
  const int lanes0 = s->list_lanes;
  auto shape_back = [&]() {
    if (s->list_lanes != lanes0) {
      s->list_lanes = lanes0;
      alloc_pblocks(s);
    }
  };
  // Once every packing still running is small, one launch runs them to the
  // end (k_fire_small): the tail of a batch then costs no launches per
  // iteration, and a small system never launches per iteration at all.  List
  // path only: its rows hold as many partners as the list kernels' (64) and
  // are chosen by a rule no wider than theirs, so the one-launch path stops
  // with SC_ERR_LIST_OVERFLOW only where the two-launch path would.
  std::vector<int> active;
  for (int k = 0; k < s->npack; ++k) active.push_back(k);
  auto all_small = [&]() {
    if (s->small_max <= 0 || s->path != SC_PATH_LIST || active.empty()) return false;
    for (int k : active)
      if (s->packs[k].n > s->small_max) return false;
    return true;
  };
  while (true) {
    if (all_small()) {
      run_small_fire(s, active, fp);
    } else if (graph) {
      if (!g.exec) g.capture([&]() { iterations(chunk); });
      g.launch();
      it += chunk;
    } else {
      for (int m = 0; m < chunk; ++m, ++it) {
        iterations(1);
        if ((it + 1) % resort_every == 0) {
          // finished packings are no longer binned by the updates: bin every
          // packing into the current buffer so the sort permutes all of them
          bin_all_sized(s, s->cur_buf, 0);
          resort(s);
        }
      }
    }
    // the scalars come back with the error word: one synchronisation per chunk
    HIPCK(hipMemcpyAsync(s->scal.data(), s->d_scal[s->cur_scal], sizeof(PackScal) * s->npack,
                         hipMemcpyDeviceToHost, s->stream));
    int e = check_errors(s);
    if (e != SC_OK) {
      s->velocity_live = false;
      s->compact = false;
      shape_back();
      return e;
    }
    active.clear();
    for (int k = 0; k < s->npack; ++k)
      if (!s->scal[k].converged) active.push_back(k);
    if (active.empty()) break;
    if (s->path == SC_PATH_LIST && 2 * (int)active.size() <= last_active) {
      if (g.exec) {  // the captured launches name the old lists
        HIPCK(hipGraphExecDestroy(g.exec));
        g.exec = nullptr;
      }
      // A tail of few particles leaves most of the GPU idle and each pass is
      // as long as one particle's chain of dependent gathers: give every
      // particle more lanes (up to 8) until the launch has about one lane
      // per SIMD lane of the chip, which shortens that chain.
      long nact = 0;
      for (int k : active) nact += s->packs[k].n;
      int lanes = s->list_lanes;
      while (lanes < 8 && nact * lanes < kTailLanes) lanes *= 2;
      if (lanes != s->list_lanes) {
        s->list_lanes = lanes;
        alloc_pblocks(s);
      }
      set_compact(s, active);
      last_active = (int)active.size();
    }
  }
  s->compact = false;
  s->velocity_live = false;
  shape_back();
  int status = SC_OK;
  for (int k = 0; k < s->npack; ++k) {
    if (iters) iters[k] = s->scal[k].iters;
    if (frms) frms[k] = s->scal[k].frms;
    if (s->scal[k].converged != 1) status = SC_ERR_NOT_CONVERGED;
  }
  // forces / energy / stress of the final state in FP64; the updates skip
  // finished packings, so the cells of every packing are filled again first
  set_prec_all(s, kFP64);
  upload_scal(s);
  s->cur_buf = 0;
  bin_all_sized(s, 0, 0);
  force_pass(s, true, 0);
  int e = check_errors(s);
  download_scal(s);
  set_prec_all(s, s->prec == SC_PREC_MIXED ? kFP32 : kFP64);
  upload_scal(s);
  HIPCK(hipStreamSynchronize(s->stream));
  if (e != SC_OK) return e;
  if (status != SC_OK) fail(s, status, "FIRE reached max_iter before F_rms < ftol");
  return status;
  API_END
}

static int export_pairs(sc_system* s, int k, int contacts, int* p, int* q, int cap, int* count) {
  const PackInfo& P = s->packs[k];
  if (s->export_cap < P.n) {
    s->dfree(s->d_export_q);
    s->dfree(s->d_export_n);
    s->d_export_q = s->dalloc<int>((size_t)P.n * kExport);
    s->d_export_n = s->dalloc<int>(P.n);
    s->export_cap = P.n;
  }
  hipLaunchKernelGGL(k_export, dim3(grid_for(P.n, 128)), dim3(128), 0, s->stream, s->B(),
                     s->d_packs, s->d_scal[s->cur_scal], s->d_cnt[s->cur_buf],
                     s->d_slot[s->cur_buf], k, contacts, s->d_export_q, s->d_export_n);
  std::vector<int> hq((size_t)P.n * kExport), hn(P.n), hid(P.n);
  HIPCK(hipMemcpyAsync(hq.data(), s->d_export_q, sizeof(int) * hq.size(), hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipMemcpyAsync(hn.data(), s->d_export_n, sizeof(int) * P.n, hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipMemcpyAsync(hid.data(), s->B().id + P.poff, sizeof(int) * P.n, hipMemcpyDeviceToHost,
                       s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  std::vector<std::pair<int, int>> pairs;
  for (int li = 0; li < P.n; ++li) {
    if (hn[li] > kExport) return fail(s, SC_ERR_LIST_OVERFLOW, "more than 64 partners in export");
    for (int j = 0; j < hn[li]; ++j) pairs.emplace_back(hid[li], hq[(size_t)li * kExport + j]);
  }
  std::sort(pairs.begin(), pairs.end());
  *count = (int)pairs.size();
  for (int i = 0; i < (int)pairs.size() && i < cap; ++i) {
    p[i] = pairs[i].first;
    q[i] = pairs[i].second;
  }
  return SC_OK;
}

int sc_get_neighbors(sc_system* s, int k, int* p, int* q, int cap, int* count) {
  if (!s || k < 0 || k >= s->npack || !count || cap < 0 || (cap > 0 && (!p || !q)))
    return SC_ERR_ARG;
  API_BEGIN
  return export_pairs(s, k, 0, p, q, cap, count);
  API_END
}

int sc_get_contacts(sc_system* s, int k, int* p, int* q, int cap, int* count) {
  if (!s || k < 0 || k >= s->npack || !count || cap < 0 || (cap > 0 && (!p || !q)))
    return SC_ERR_ARG;
  API_BEGIN
  return export_pairs(s, k, 1, p, q, cap, count);
  API_END
}

int sc_timing_enable(sc_system* s, int enable) {
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  resolve_events(s);
  s->timing = enable != 0;
  return SC_OK;
  API_END
}

int sc_timing_get(sc_system* s, int kernel, double* ms_total, long long* launches) {
  if (!s || kernel < 0 || kernel >= SC_KERNEL_COUNT) return SC_ERR_ARG;
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  resolve_events(s);
  if (ms_total) *ms_total = s->kt_ms[kernel];
  if (launches) *launches = s->kt_n[kernel];
  return SC_OK;
  API_END
}

int sc_timing_reset(sc_system* s) {
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  resolve_events(s);
  for (int i = 0; i < SC_KERNEL_COUNT; ++i) {
    s->kt_ms[i] = 0;
    s->kt_n[i] = 0;
  }
  return SC_OK;
  API_END
}

int sc_synchronize(sc_system* s) {
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  return check_errors(s);
  API_END
}

// Project autotuner: times the force pass of the system as it is (no state
// change beyond the forces) for each candidate path and launch shape, keeps the
// fastest.  reps passes per candidate after one warm-up pass.
int sc_set_pad(sc_system* s, int k, double pad) {
  if (!s || k < -1 || k >= s->npack || !(pad >= 0) || !std::isfinite(pad)) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  for (int j = 0; j < s->npack; ++j) {
    if (k >= 0 && j != k) continue;
    PackInfo& P = s->packs[j];
    double p = pad;
    cell_grid(s->scal[j].Lx, s->scal[j].Ly, s->Rmax[j], s->Amax[j], p, P.rows, P.cols);
    P.pad = p;
  }
  build_tiles(s);
  alloc_grid(s);
  s->cur_buf = 0;
  bin_all_sized(s, 0, 0);
  resort(s);  // also rebuilds every row with the new skin
  return check_errors(s);
  API_END
}

// Cell-size autotuner: each candidate padding (all packings alike) runs
// `steps` Heun steps of the system's own dynamics from the same saved state,
// timed on the host around a synchronised stream (row rebuilds included, which
// is where the padding pays or costs); the fastest is kept and the state is
// put back.  Positions, angles and strain are restored exactly; forces are
// those of the restored state's next pass.
int sc_autotune_pad(sc_system* s, int ncand, const double* pads, int steps, double* ms_out,
                    double* chosen) {
  if (!s || ncand < 1 || !pads || steps < 1) return SC_ERR_ARG;
  for (int c = 0; c < ncand; ++c)
    if (!(pads[c] >= 0) || !std::isfinite(pads[c])) return SC_ERR_ARG;
  API_BEGIN
  std::vector<double> x(s->ntot), y(s->ntot), phi(s->ntot);
  int e = sc_get_positions(s, x.data(), y.data(), phi.data());
  if (e != SC_OK) return e;
  download_scal(s);
  std::vector<double> g0(s->npack), gt0(s->npack);
  for (int k = 0; k < s->npack; ++k) {
    g0[k] = s->scal[k].gamma;
    gt0[k] = s->scal[k].gamma_total;
  }
  auto restore = [&]() -> int {
    int st = sc_set_positions(s, x.data(), y.data(), phi.data());
    for (int k = 0; k < s->npack && st == SC_OK; ++k) st = sc_set_gamma(s, k, g0[k], gt0[k]);
    return st;
  };
  int best = 0;
  double best_ms = 1e300;
  for (int c = 0; c < ncand; ++c) {
    if ((e = sc_set_pad(s, -1, pads[c])) != SC_OK || (e = restore()) != SC_OK) return e;
    HIPCK(hipStreamSynchronize(s->stream));
    const auto t0 = std::chrono::steady_clock::now();
    if ((e = sc_strain_steps(s, steps, 0, nullptr)) != SC_OK) {
      // the trial failed: put the caller's state back, keep the trial's message
      const std::string msg = s->err;
      (void)restore();
      s->err = msg;
      return e;
    }
    HIPCK(hipStreamSynchronize(s->stream));
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (ms_out) ms_out[c] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  if ((e = sc_set_pad(s, -1, pads[best])) != SC_OK || (e = restore()) != SC_OK) return e;
  if (chosen) *chosen = pads[best];
  return SC_OK;
  API_END
}

int sc_autotune(sc_system* s, int reps, int* path_out, int* tile_cols_out, int* block_out,
                double* us_out) {
  if (!s || reps < 1) return SC_ERR_ARG;
  API_BEGIN
  struct Cand {
    int path, tile, block, lanes;
  };
  std::vector<Cand> cands;
  for (int lanes : {1, 2, 4, 8})
    for (int blk : {64, 128, 256}) cands.push_back({SC_PATH_LIST, 0, blk, lanes});
  for (Cand c : {Cand{SC_PATH_TILE, 8, 128, 0}, Cand{SC_PATH_TILE, 12, 128, 0},
                 Cand{SC_PATH_TILE, 16, 256, 0}})
    cands.push_back(c);
  hipEvent_t e0, e1;
  HIPCK(hipEventCreate(&e0));
  HIPCK(hipEventCreate(&e1));
  const bool was_timing = s->timing;
  s->timing = false;
  // Three interleaved rounds over the candidates, each candidate scored by its
  // fastest round: clock and cache drift between candidates averages out.
  std::vector<double> score(cands.size(), 1e30);
  for (int round = 0; round < 3; ++round) {
    for (size_t ci = 0; ci < cands.size(); ++ci) {
      const Cand& c = cands[ci];
      if (c.path == SC_PATH_TILE) {
        if (tile_lds_bytes(3 * (c.tile + 2) * s->cap, c.block / kLanes, c.block) > (size_t)s->lds_max)
          continue;
        if (sc_set_tiling(s, c.tile, c.block) != SC_OK) continue;
      } else if (sc_set_list_block(s, c.block, c.lanes) != SC_OK) {
        continue;
      }
      sc_set_path(s, c.path);
      // The Heun predictor is the representative pass: same gathers and pair
      // math as the corrector, and it leaves the state (positions) untouched.
      // Its output buckets are zeroed before every repetition (same cost for
      // every candidate, so the ranking is unaffected).
      PassSpec ps;
      ps.mode = kModePred;
      ps.read_buf = s->cur_buf;
      ps.fill_buf = (s->cur_buf + 1) % 3;
      ps.clear_buf = (s->cur_buf + 2) % 3;
      auto one = [&]() {
        HIPCK(hipMemsetAsync(s->d_cnt[ps.fill_buf], 0, sizeof(int) * s->ncells, s->stream));
        launch_force(s, ps);
      };
      PassSpec fwd;
      fwd.mode = kModeForces;
      fwd.read_buf = s->cur_buf;
      launch_force(s, fwd);  // warm-up: rebuilds the lists at the current state
      one();
      HIPCK(hipEventRecord(e0, s->stream));
      for (int r = 0; r < reps; ++r) one();
      HIPCK(hipEventRecord(e1, s->stream));
      HIPCK(hipEventSynchronize(e1));
      float ms = 0;
      HIPCK(hipEventElapsedTime(&ms, e0, e1));
      score[ci] = std::min(score[ci], 1e3 * ms / reps);
    }
  }
  size_t bi = 0;
  for (size_t ci = 1; ci < cands.size(); ++ci)
    if (score[ci] < score[bi]) bi = ci;
  const double best = score[bi];
  const Cand pick = cands[bi];
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  s->timing = was_timing;
  if (pick.path == SC_PATH_TILE) sc_set_tiling(s, pick.tile, pick.block);
  else sc_set_list_block(s, pick.block, pick.lanes);
  sc_set_path(s, pick.path);
  int e = check_errors(s);
  if (e != SC_OK) return e;
  if (path_out) *path_out = pick.path;
  if (tile_cols_out) *tile_cols_out = pick.path == SC_PATH_TILE ? pick.tile : pick.lanes;
  if (block_out) *block_out = pick.block;
  if (us_out) *us_out = best;
  return SC_OK;
  API_END
}

int sc_autotune_fire(sc_system* s, const double* params, double switch_frms, int iterations,
                     int* path_out, int* shape_out, int* block_out, double* ms_out) {
  if (!s || !params || iterations < 1) return SC_ERR_ARG;
  API_BEGIN
  struct Cand {
    int path, tile, block, lanes;
  };
  std::vector<Cand> cands;
  for (int lanes : {1, 2, 4, 8})
    for (int blk : {64, 128, 256}) cands.push_back({SC_PATH_LIST, 0, blk, lanes});
  for (Cand c : {Cand{SC_PATH_TILE, 8, 128, 0}, Cand{SC_PATH_TILE, 12, 128, 0},
                 Cand{SC_PATH_TILE, 16, 256, 0}})
    cands.push_back(c);
  std::vector<double> x(s->ntot), y(s->ntot), phi(s->ntot);
  int e = sc_get_positions(s, x.data(), y.data(), phi.data());
  if (e != SC_OK) return e;
  std::vector<double> p(params, params + 10);
  p[9] = iterations;  // max_iter
  // Two interleaved rounds, each candidate scored by its faster one.  Every
  // trial starts from the saved positions (FIRE zeroes the velocities and
  // leaves the box alone), so the candidates run the same minimisation.
  std::vector<double> score(cands.size(), 1e300);
  for (int round = 0; round < 2; ++round) {
    for (size_t ci = 0; ci < cands.size(); ++ci) {
      const Cand& c = cands[ci];
      if (c.path == SC_PATH_TILE) {
        if (tile_lds_bytes(3 * (c.tile + 2) * s->cap, c.block / kLanes, c.block) > (size_t)s->lds_max)
          continue;
        if (sc_set_tiling(s, c.tile, c.block) != SC_OK) continue;
      } else if (sc_set_list_block(s, c.block, c.lanes) != SC_OK) {
        continue;
      }
      if ((e = sc_set_path(s, c.path)) != SC_OK) return e;
      if ((e = sc_set_positions(s, x.data(), y.data(), phi.data())) != SC_OK) return e;
      HIPCK(hipStreamSynchronize(s->stream));
      const auto t0 = std::chrono::steady_clock::now();
      e = sc_fire_minimize(s, p.data(), switch_frms, nullptr, nullptr);
      HIPCK(hipStreamSynchronize(s->stream));
      if (e != SC_OK && e != SC_ERR_NOT_CONVERGED) {
        // the trial failed (e.g. a non-finite coordinate): put the caller's
        // positions back, keep the trial's message, report its status
        const std::string msg = s->err;
        (void)sc_set_positions(s, x.data(), y.data(), phi.data());
        s->err = msg;
        return e;
      }
      const double ms =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      score[ci] = std::min(score[ci], ms);
    }
  }
  size_t bi = 0;
  for (size_t ci = 1; ci < cands.size(); ++ci)
    if (score[ci] < score[bi]) bi = ci;
  const Cand pick = cands[bi];
  if (pick.path == SC_PATH_TILE) sc_set_tiling(s, pick.tile, pick.block);
  else sc_set_list_block(s, pick.block, pick.lanes);
  if ((e = sc_set_path(s, pick.path)) != SC_OK) return e;
  if ((e = sc_set_positions(s, x.data(), y.data(), phi.data())) != SC_OK) return e;
  s->err.clear();
  if (path_out) *path_out = pick.path;
  if (shape_out) *shape_out = pick.path == SC_PATH_TILE ? pick.tile : pick.lanes;
  if (block_out) *block_out = pick.block;
  if (ms_out) *ms_out = score[bi];
  return SC_OK;
  API_END
}

}  // extern "C"

namespace sc {
int system_view(sc_system* s, SystemView* v) {
  if (!s || !v) return SC_ERR_ARG;
  v->device = s->device;
  v->stream = s->stream;
  v->ntot = s->ntot;
  v->npack = s->npack;
  v->ncells = s->ncells;
  v->b = s->B();
  v->packs = s->d_packs;
  v->scal = s->d_scal[s->cur_scal];
  v->cnt = s->d_cnt[s->cur_buf];
  v->slot = s->d_slot[s->cur_buf];
  v->pack_of = s->d_pack_of;
  v->pack_sum = s->d_pack_sum;
  return SC_OK;
}

int system_energy_pass(sc_system* s) {
  API_BEGIN
  s->cur_buf = 0;
  bin_all(s, 0, 0);
  segvec(s);
  force_rebuild(s);
  force_pass(s, true, 0);
  return SC_OK;
  API_END
}

int system_positions_changed(sc_system* s) {
  if (!s) return SC_ERR_ARG;
  API_BEGIN
  s->cur_buf = 0;
  bin_all_sized(s, 0, 1);
  segvec(s);
  resort(s);
  return check_errors(s);
  API_END
}
}  // namespace sc



namespace {
// reconfigure_cells (src/cpp/resize_dynamics.cpp:106-121): when the box length
// changes the number of cell rows or columns, rebuild the grid, tiles and
// buckets and re-sort.
int reconfigure_if_needed(sc_system* s, const std::vector<std::pair<double, double>>* Lhint) {
  if (Lhint) {
    bool any = false;
    for (int k = 0; k < s->npack; ++k) {
      double pad = s->packs[k].pad;
      int rows, cols;
      cell_grid((*Lhint)[k].first, (*Lhint)[k].second, s->Rmax[k], s->Amax[k], pad, rows, cols);
      any |= rows != s->packs[k].rows || cols != s->packs[k].cols;
    }
    if (!any) return SC_OK;
  }
  download_scal(s);  // device L is authoritative
  if (!regrid(s)) return SC_OK;
  s->cur_buf = 0;
  bin_all_sized(s, 0, 0);
  resort(s);
  return check_errors(s);
}

// The cell grid of every packing for its current box (host copy of the
// scalars); reallocates the grid buffers when a grid changed.  Returns
// whether one did.  The caller bins and re-sorts.
bool regrid(sc_system* s) {
  bool changed = false;
  for (int k = 0; k < s->npack; ++k) {
    PackInfo& P = s->packs[k];
    double pad = P.pad;
    int rows, cols;
    cell_grid(s->scal[k].Lx, s->scal[k].Ly, s->Rmax[k], s->Amax[k], pad, rows, cols);
    if (rows != P.rows || cols != P.cols) {
      P.rows = rows;
      P.cols = cols;
      changed = true;
    }
  }
  if (!changed) return false;
  build_tiles(s);
  alloc_grid(s);
  return true;
}
}  // namespace
