// Host side of the C ABI (include/spherocyls.h): system lifecycle, cell grids
// and tiles, kernel orchestration on the system's own HIP stream, timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spherocyls.h"
#include "sc_internal.h"

namespace sc {
// kernels (sc_device.hip)
template <int MODE>
__global__ void k_tile(TileArgs a);
__global__ void k_bin_all(Bank b, const int* pack_of, const PackInfo* packs, const PackScal* scal,
                          int* cnt, int* slot, int* err, int ntot, int wrap_angles);
__global__ void k_update(Bank b, const int* pack_of, const PackInfo* packs, const PackScal* scal,
                         int* cnt, int* slot, int* cnt_clear, int ncells_total, int* err, int ntot,
                         int mode, double max_move);
__global__ void k_affine(Bank b, const int* pack_of, const double* scale, int ntot);
__global__ void k_scan_cells(const PackInfo* packs, const int* cnt, int* cstart);
__global__ void k_sort_cells(const int* cstart, int* cnt, int* slot, const int* id, int* perm,
                             int ncells_total);
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

#define HIPCK(call)                                 \
  do {                                              \
    hipError_t e_ = (call);                         \
    if (e_ != hipSuccess) throw HipError{e_, #call}; \
  } while (0)

struct StatusError {
  int status;
  std::string msg;
};

// Cell grid of configure_cells (src/cpp/spherocyl_box.cpp:228-241).
void cell_grid(double L, double Rmax, double Amax, double& pad, int& rows, int& cols) {
  double wmin = 2.24 * (Rmax + Amax) + pad;
  double hmin = 2 * (Rmax + Amax) + pad;
  rows = std::max(static_cast<int>(L / hmin), 1);
  cols = std::max(static_cast<int>(L / wmin), 1);
  if (pad == 0) pad = std::max(L / cols - wmin, L / rows - hmin);
}

struct EventPair {
  hipEvent_t a, b;
  int kernel;
};

}  // namespace

struct sc_system {
  int device = 0;
  hipStream_t stream = nullptr;
  int npack = 0, ntot = 0;
  std::vector<int> offsets;
  std::vector<PackInfo> packs;
  std::vector<PackScal> scal;  // host mirror of d_scal[cur_scal] (valid after sync_scal)
  std::vector<double> Rmax, Amax, packing, filled;
  int tile_cols = 8, block = 128;
  int lds_max = 64 * 1024;  // LDS bytes a workgroup may use on this device
  int prec = SC_PREC_FP64;
  int ncells = 0, ntiles = 0;
  std::vector<int4> tiles;

  PackInfo* d_packs = nullptr;
  PackScal* d_scal[2] = {nullptr, nullptr};
  int cur_scal = 0;
  int4* d_tiles = nullptr;
  int* d_cnt[3] = {nullptr, nullptr, nullptr};
  int* d_slot[3] = {nullptr, nullptr, nullptr};
  int cur_buf = 0;
  int* d_cstart = nullptr;
  int* d_perm = nullptr;
  Bank bank[2]{};
  int cur_bank = 0;
  double *d_px = nullptr, *d_py = nullptr, *d_pphi = nullptr;
  double *d_v0x = nullptr, *d_v0y = nullptr, *d_v0t = nullptr;
  int* d_pack_of = nullptr;
  double* d_tile_part = nullptr;
  double* d_pack_sum = nullptr;
  unsigned* d_pack_done = nullptr;
  double* d_se_ring = nullptr;
  size_t se_ring_cap = 0;
  int* d_err = nullptr;
  double* d_io = nullptr;  // 3 * ntot staging for caller-order transfers
  int* d_export_q = nullptr;
  int* d_export_n = nullptr;
  int export_cap = 0;
  double* d_scale = nullptr;
  bool velocity_live = false;

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
  int coff = 0;
  for (int k = 0; k < s->npack; ++k) {
    PackInfo& P = s->packs[k];
    P.coff = coff;
    coff += P.rows * P.cols;
    int tw = P.cols >= 3 ? std::min(s->tile_cols, P.cols - 2) : P.cols;
    tw = std::max(1, std::min(tw, kMaxTileCols));
    P.tile_cols = tw;
    std::vector<int4> spatial;
    for (int r = 0; r < P.rows; ++r)
      for (int c0 = 0; c0 < P.cols; c0 += tw) spatial.push_back(make_int4(k, r, c0, std::min(tw, P.cols - c0)));
    P.toff = (int)s->tiles.size();
    P.ntiles = (int)spatial.size();
    // XCD-aware order: workgroup b runs on XCD b % 8, so give each XCD a
    // spatially contiguous run of tiles (neighbouring tiles share staged rows).
    const int nx = 8, T = P.ntiles;
    std::vector<int> count(nx, 0);
    for (int i = 0; i < T; ++i) count[(P.toff + i) % nx]++;
    std::vector<int> start(nx, 0);
    for (int x = 1; x < nx; ++x) start[x] = start[x - 1] + count[x - 1];
    std::vector<int> used(nx, 0);
    std::vector<int4> launch(T);
    for (int i = 0; i < T; ++i) {
      int x = (P.toff + i) % nx;
      launch[i] = spatial[start[x] + used[x]++];
    }
    s->tiles.insert(s->tiles.end(), launch.begin(), launch.end());
  }
  s->ncells = coff;
  s->ntiles = (int)s->tiles.size();
}

// (Re)allocate everything whose size depends on the cell grid / tiles.
void alloc_grid(sc_system* s) {
  for (int i = 0; i < 3; ++i) {
    s->dfree(s->d_cnt[i]);
    s->dfree(s->d_slot[i]);
    s->d_cnt[i] = s->dalloc<int>(s->ncells);
    s->d_slot[i] = s->dalloc<int>((size_t)s->ncells * kCap);
    HIPCK(hipMemsetAsync(s->d_cnt[i], 0, sizeof(int) * s->ncells, s->stream));
  }
  s->dfree(s->d_cstart);
  s->d_cstart = s->dalloc<int>(s->ncells);
  s->dfree(s->d_tiles);
  s->d_tiles = s->dalloc<int4>(s->ntiles);
  HIPCK(hipMemcpyAsync(s->d_tiles, s->tiles.data(), sizeof(int4) * s->ntiles, hipMemcpyHostToDevice,
                       s->stream));
  s->dfree(s->d_tile_part);
  s->d_tile_part = s->dalloc<double>((size_t)s->ntiles * kNSum);
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
}

Bank alloc_bank(sc_system* s) {
  Bank b{};
  size_t n = s->ntot;
  b.x = s->dalloc<double>(n);
  b.y = s->dalloc<double>(n);
  b.phi = s->dalloc<double>(n);
  b.R = s->dalloc<double>(n);
  b.A = s->dalloc<double>(n);
  b.I = s->dalloc<double>(n);
  b.RC = s->dalloc<double>(n);
  b.fx = s->dalloc<double>(n);
  b.fy = s->dalloc<double>(n);
  b.tau = s->dalloc<double>(n);
  b.vx = s->dalloc<double>(n);
  b.vy = s->dalloc<double>(n);
  b.vpsi = s->dalloc<double>(n);
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
    return fail(s, SC_ERR_CELL_OVERFLOW, "a cell holds more particles than its bucket (16)");
  if (e & kErrList)
    return fail(s, SC_ERR_LIST_OVERFLOW, "a particle has more near candidates than its list holds");
  return fail(s, SC_ERR_ARG, "non-finite particle coordinate");
}

int grid_for(int n, int block) { return (n + block - 1) / block; }

// Bin every particle of the current bank into bucket buffer `buf` (cleared).
void bin_all(sc_system* s, int buf, int wrap_angles) {
  HIPCK(hipMemsetAsync(s->d_cnt[buf], 0, sizeof(int) * s->ncells, s->stream));
  hipLaunchKernelGGL(k_bin_all, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_packs, s->d_scal[s->cur_scal], s->d_cnt[buf],
                     s->d_slot[buf], s->d_err, s->ntot, wrap_angles);
}

// Counting sort of the current bank by the cells of bucket buffer cur_buf.
void resort(sc_system* s) {
  KernelTimer t(s, SC_KERNEL_SORT);
  int buf = s->cur_buf;
  hipLaunchKernelGGL(k_scan_cells, dim3(s->npack), dim3(256), 0, s->stream, s->d_packs,
                     s->d_cnt[buf], s->d_cstart);
  hipLaunchKernelGGL(k_sort_cells, dim3(grid_for(s->ncells, 256)), dim3(256), 0, s->stream,
                     s->d_cstart, s->d_cnt[buf], s->d_slot[buf], s->B().id, s->d_perm, s->ncells);
  Bank& src = s->bank[s->cur_bank];
  Bank& dst = s->bank[1 - s->cur_bank];
  hipLaunchKernelGGL(k_gather, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, src, dst,
                     s->d_perm, s->ntot, s->velocity_live ? 1 : 0);
  s->cur_bank = 1 - s->cur_bank;
}

TileArgs base_args(sc_system* s) {
  TileArgs a{};
  a.packs = s->d_packs;
  a.scal = s->d_scal[s->cur_scal];
  a.scal_next = s->d_scal[1 - s->cur_scal];
  a.tiles = s->d_tiles;
  a.cnt = s->d_cnt[s->cur_buf];
  a.slot = s->d_slot[s->cur_buf];
  a.ncells_total = s->ncells;
  a.b = s->B();
  a.sx = a.b.x;
  a.sy = a.b.y;
  a.sphi = a.b.phi;
  a.px = s->d_px;
  a.py = s->d_py;
  a.pphi = s->d_pphi;
  a.v0x = s->d_v0x;
  a.v0y = s->d_v0y;
  a.v0t = s->d_v0t;
  a.tile_part = s->d_tile_part;
  a.pack_sum = s->d_pack_sum;
  a.pack_done = s->d_pack_done;
  a.se_ring = s->d_se_ring;
  a.err = s->d_err;
  a.npack = s->npack;
  a.list_groups = s->block / kLanes;
  a.stage_cap = 3 * (s->tile_cols + 2) * kCap;
  return a;
}

template <int MODE>
void launch_tile(sc_system* s, const TileArgs& a, int kernel_id) {
  size_t lds = tile_lds_bytes(a.stage_cap, a.list_groups, s->block);
  if (lds > (size_t)s->lds_max) throw StatusError{SC_ERR_ARG, "tile too wide for LDS"};
  if (lds > 64 * 1024) {  // wide tiles need more than the default dynamic LDS
    HIPCK(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_tile<MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  KernelTimer t(s, kernel_id);
  hipLaunchKernelGGL(k_tile<MODE>, dim3(s->ntiles), dim3(s->block), lds, s->stream, a);
}

void ensure_se_ring(sc_system* s, size_t nrec) {
  size_t need = nrec * s->npack * kNSe;
  if (need <= s->se_ring_cap) return;
  s->dfree(s->d_se_ring);
  s->d_se_ring = s->dalloc<double>(need);
  s->se_ring_cap = need;
}

// Plain force pass at the current state (calc_forces / calc_se).
void force_pass(sc_system* s, bool se, int se_every) {
  TileArgs a = base_args(s);
  a.se_always = se ? 1 : 0;
  a.se_every = se_every;
  a.do_resid = 1;
  a.write_forces = 1;
  launch_tile<kModeForces>(s, a, SC_KERNEL_FORCE);
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

int reconfigure_if_needed(sc_system* s, const std::vector<double>* Lhint);

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int sc_version(void) { return 10000; }

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

  sc_system* s = new sc_system();
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
    s->packing.assign(npack, 0.0);
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
      s->packing[k] = filled / (L[k] * L[k]);
      s->filled[k] = filled;
      P.pad = pad[k];
      cell_grid(L[k], s->Rmax[k], s->Amax[k], P.pad, P.rows, P.cols);
      PackScal& S = s->scal[k];
      S.Lx = S.Ly = L[k];
      S.prec = kFP64;
    }
    build_tiles(s);

    s->d_packs = s->dalloc<PackInfo>(npack);
    s->d_scal[0] = s->dalloc<PackScal>(npack);
    s->d_scal[1] = s->dalloc<PackScal>(npack);
    s->d_pack_sum = s->dalloc<double>((size_t)npack * kNSum);
    s->d_pack_done = s->dalloc<unsigned>(npack);
    HIPCK(hipMemsetAsync(s->d_pack_done, 0, sizeof(unsigned) * npack, s->stream));
    HIPCK(hipMemsetAsync(s->d_pack_sum, 0, sizeof(double) * npack * kNSum, s->stream));
    s->d_err = s->dalloc<int>(1);
    HIPCK(hipMemsetAsync(s->d_err, 0, sizeof(int), s->stream));
    s->bank[0] = alloc_bank(s);
    s->bank[1] = alloc_bank(s);
    s->d_px = s->dalloc<double>(ntot);
    s->d_py = s->dalloc<double>(ntot);
    s->d_pphi = s->dalloc<double>(ntot);
    s->d_v0x = s->dalloc<double>(ntot);
    s->d_v0y = s->dalloc<double>(ntot);
    s->d_v0t = s->dalloc<double>(ntot);
    s->d_pack_of = s->dalloc<int>(ntot);
    s->d_perm = s->dalloc<int>(ntot);
    HIPCK(hipMemsetAsync(s->d_perm, 0, sizeof(int) * ntot, s->stream));
    s->d_io = s->dalloc<double>((size_t)3 * ntot);
    s->d_scale = s->dalloc<double>(npack);
    alloc_grid(s);
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
    bin_all(s, 0, 0);
    resort(s);
    HIPCK(hipStreamSynchronize(s->stream));
    int st = check_errors(s);
    if (st != SC_OK) {
      std::string m = s->err;
      sc_system_destroy(s);
      return st;
    }
  } catch (const HipError& e) {
    std::fprintf(stderr, "spherocyls: %s failed: %s\n", e.what, hipGetErrorString(e.e));
    sc_system_destroy(s);
    return SC_ERR_HIP;
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

#define API_BEGIN                      \
  if (!s) return SC_ERR_ARG;           \
  try {                                \
    HIPCK(hipSetDevice(s->device));
#define API_END                                                                   \
  }                                                                               \
  catch (const HipError& e) {                                                     \
    return fail(s, SC_ERR_HIP, std::string(e.what) + ": " + hipGetErrorString(e.e)); \
  }                                                                               \
  catch (const StatusError& e) {                                                  \
    return fail(s, e.status, e.msg);                                              \
  }

int sc_set_dynamics(sc_system* s, int k, double strain_rate, double resize_rate, double step) {
  if (!s || k < -1 || k >= s->npack || !(step > 0)) return SC_ERR_ARG;
  API_BEGIN
  for (int i = 0; i < s->npack; ++i) {
    if (k >= 0 && i != k) continue;
    s->packs[i].strain_rate = strain_rate;
    s->packs[i].resize_rate = resize_rate;
    s->packs[i].dt = step;
  }
  HIPCK(hipMemcpyAsync(s->d_packs, s->packs.data(), sizeof(PackInfo) * s->npack,
                       hipMemcpyHostToDevice, s->stream));
  HIPCK(hipStreamSynchronize(s->stream));
  return SC_OK;
  API_END
}

int sc_set_gamma(sc_system* s, int k, double gamma, double gamma_total) {
  if (!s || k < -1 || k >= s->npack) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  for (int i = 0; i < s->npack; ++i) {
    if (k >= 0 && i != k) continue;
    s->scal[i].gamma = gamma;
    s->scal[i].gamma_total = gamma_total;
  }
  upload_scal(s);
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

int sc_set_tiling(sc_system* s, int tile_cols, int block_threads) {
  if (!s || tile_cols < 1 || tile_cols > kMaxTileCols) return SC_ERR_ARG;
  if (block_threads != 64 && block_threads != 128 && block_threads != 256) return SC_ERR_ARG;
  if (tile_lds_bytes(3 * (tile_cols + 2) * kCap, block_threads / kLanes, block_threads) >
      (size_t)s->lds_max)
    return fail(s, SC_ERR_ARG, "tile too wide for the workgroup's LDS");
  API_BEGIN
  HIPCK(hipStreamSynchronize(s->stream));
  s->tile_cols = tile_cols;
  s->block = block_threads;
  build_tiles(s);
  s->dfree(s->d_tiles);
  s->d_tiles = s->dalloc<int4>(s->ntiles);
  HIPCK(hipMemcpyAsync(s->d_tiles, s->tiles.data(), sizeof(int4) * s->ntiles, hipMemcpyHostToDevice,
                       s->stream));
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
  if (tile_cols) *tile_cols = s->tile_cols;
  if (block_threads) *block_threads = s->block;
  return SC_OK;
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
  bin_all(s, 0, 1);
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

// One Heun sub-step (shear or resize) on every packing.  Bucket buffers
// rotate: c holds the cells of x, the predictor bins x~ into c+1 and zeroes
// c+2, the corrector reads c+1, bins x' into c+2 and zeroes c (ready to take
// the next step's predicted cells).
static void heun_substep(sc_system* s, bool resize, bool shrink, bool corrected, int se_every,
                         double scale) {
  const int c0 = s->cur_buf, c1 = (c0 + 1) % 3, c2 = (c0 + 2) % 3;
  TileArgs a = base_args(s);
  a.resize = resize;
  a.shrink = shrink;
  a.corrected_image = corrected;
  a.resize_scale = scale;
  a.se_every = se_every;
  a.cnt_fill = s->d_cnt[c1];
  a.slot_fill = s->d_slot[c1];
  a.cnt_clear = s->d_cnt[c2];
  launch_tile<kModePred>(s, a, SC_KERNEL_PREDICT);
  TileArgs c = base_args(s);
  c.resize = resize;
  c.shrink = shrink;
  c.corrected_image = corrected;
  c.resize_scale = scale;
  c.cnt = s->d_cnt[c1];
  c.slot = s->d_slot[c1];
  c.sx = s->d_px;
  c.sy = s->d_py;
  c.sphi = s->d_pphi;
  c.cnt_fill = s->d_cnt[c2];
  c.slot_fill = s->d_slot[c2];
  c.cnt_clear = s->d_cnt[c0];
  launch_tile<kModeCorr>(s, c, SC_KERNEL_CORRECT);
  s->cur_buf = c2;
  s->cur_scal = 1 - s->cur_scal;
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
  const int resort_every = 1000;
  for (int st = 0; st < nsteps; ++st) {
    heun_substep(s, false, false, false, se_every, 1.0);
    if ((st + 1) % resort_every == 0 && st + 1 < nsteps) resort(s);
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
  std::vector<double> Lm(s->npack);
  for (int k = 0; k < s->npack; ++k) Lm[k] = s->scal[k].Lx;
  for (int st = 0; st < nsteps; ++st) {
    // all packings share one resize rate and step in this call
    double scale = std::pow(1.0 - s->packs[0].resize_rate, s->packs[0].dt);
    heun_substep(s, true, shrink != 0, corrected_image != 0, se_every, scale);
    for (auto& v : Lm) v = shrink ? v * scale : v / scale;
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

int sc_gd_steps(sc_system* s, int nmax, double emin, int* steps_done) {
  if (!s || nmax < 0) return SC_ERR_ARG;
  API_BEGIN
  reset_step_counters(s);
  clear_other_buffers(s);
  FireParams fp{};
  fp.emin = emin;
  fp.max_iter = nmax;
  const int chunk = 64;
  int done = 0;
  while (true) {
    for (int m = 0; m < chunk; ++m) {
      TileArgs a = base_args(s);
      a.do_gd = 1;
      a.se_always = 1;
      a.write_forces = 1;
      a.fire = fp;
      launch_tile<kModeForces>(s, a, SC_KERNEL_FIRE_FORCE);
      int b1 = (s->cur_buf + 1) % 3, b2 = (s->cur_buf + 2) % 3;
      {
        KernelTimer t(s, SC_KERNEL_FIRE_UPDATE);
        hipLaunchKernelGGL(k_update, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                           s->d_pack_of, s->d_packs, s->d_scal[1 - s->cur_scal], s->d_cnt[b1],
                           s->d_slot[b1], s->d_cnt[b2], s->ncells, s->d_err, s->ntot, 0, 0.0);
      }
      s->cur_buf = b1;
      s->cur_scal = 1 - s->cur_scal;
    }
    int e = check_errors(s);
    if (e != SC_OK) return e;
    download_scal(s);
    bool all = true;
    for (auto& S : s->scal) all &= S.converged != 0;
    if (all) break;
  }
  for (auto& S : s->scal) done = std::max(done, S.iters);
  if (steps_done) *steps_done = done;
  // positions are final; leave forces, energy and stress of the final state
  force_pass(s, true, 0);
  return check_errors(s);
  API_END
}

int sc_affine_resize(sc_system* s, const double* scale) {
  if (!s || !scale) return SC_ERR_ARG;
  for (int k = 0; k < s->npack; ++k)
    if (!(scale[k] > 0)) return SC_ERR_ARG;
  API_BEGIN
  download_scal(s);
  for (int k = 0; k < s->npack; ++k) {
    s->scal[k].Lx *= scale[k];
    s->scal[k].Ly *= scale[k];
  }
  upload_scal(s);
  HIPCK(hipMemcpyAsync(s->d_scale, scale, sizeof(double) * s->npack, hipMemcpyHostToDevice,
                       s->stream));
  hipLaunchKernelGGL(k_affine, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                     s->d_pack_of, s->d_scale, s->ntot);
  int e = reconfigure_if_needed(s, nullptr);
  if (e != SC_OK) return e;
  s->cur_buf = 0;
  bin_all(s, 0, 0);
  resort(s);
  return check_errors(s);
  API_END
}

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
  const int chunk = 32;
  const int resort_every = 256;
  int it = 0;
  while (true) {
    for (int m = 0; m < chunk; ++m, ++it) {
      TileArgs a = base_args(s);
      a.do_fire = 1;
      a.write_forces = 1;
      a.fire = fp;
      launch_tile<kModeForces>(s, a, SC_KERNEL_FIRE_FORCE);
      int b1 = (s->cur_buf + 1) % 3, b2 = (s->cur_buf + 2) % 3;
      {
        KernelTimer t(s, SC_KERNEL_FIRE_UPDATE);
        hipLaunchKernelGGL(k_update, dim3(grid_for(s->ntot, 256)), dim3(256), 0, s->stream, s->B(),
                           s->d_pack_of, s->d_packs, s->d_scal[1 - s->cur_scal], s->d_cnt[b1],
                           s->d_slot[b1], s->d_cnt[b2], s->ncells, s->d_err, s->ntot, 1,
                           fp.max_move);
      }
      s->cur_buf = b1;
      s->cur_scal = 1 - s->cur_scal;
      if ((it + 1) % resort_every == 0) resort(s);
    }
    int e = check_errors(s);
    if (e != SC_OK) {
      s->velocity_live = false;
      return e;
    }
    download_scal(s);
    bool all = true;
    for (auto& S : s->scal) all &= S.converged != 0;
    if (all) break;
  }
  s->velocity_live = false;
  int status = SC_OK;
  for (int k = 0; k < s->npack; ++k) {
    if (iters) iters[k] = s->scal[k].iters;
    if (frms) frms[k] = s->scal[k].frms;
    if (s->scal[k].converged != 1) status = SC_ERR_NOT_CONVERGED;
  }
  // forces / energy / stress of the final state in FP64
  download_scal(s);
  set_prec_all(s, kFP64);
  upload_scal(s);
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

}  // extern "C"

namespace {
// reconfigure_cells (src/cpp/resize_dynamics.cpp:106-121): when the box length
// changes the number of cell rows or columns, rebuild the grid, tiles and
// buckets and re-sort.  Uses the host mirror of L in s->scal.
int reconfigure_if_needed(sc_system* s, const std::vector<double>* Lhint) {
  if (Lhint) {
    bool any = false;
    for (int k = 0; k < s->npack; ++k) {
      double pad = s->packs[k].pad;
      int rows, cols;
      cell_grid((*Lhint)[k], s->Rmax[k], s->Amax[k], pad, rows, cols);
      any |= rows != s->packs[k].rows || cols != s->packs[k].cols;
    }
    if (!any) return SC_OK;
  }
  download_scal(s);  // device L is authoritative
  bool changed = false;
  for (int k = 0; k < s->npack; ++k) {
    PackInfo& P = s->packs[k];
    double pad = P.pad;
    int rows, cols;
    cell_grid(s->scal[k].Lx, s->Rmax[k], s->Amax[k], pad, rows, cols);
    if (rows != P.rows || cols != P.cols) {
      P.rows = rows;
      P.cols = cols;
      changed = true;
    }
  }
  if (!changed) return SC_OK;
  build_tiles(s);
  alloc_grid(s);
  s->cur_buf = 0;
  bin_all(s, 0, 0);
  resort(s);
  return check_errors(s);
}
}  // namespace
