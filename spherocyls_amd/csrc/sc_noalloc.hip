// sc_calc_se_noalloc: the calc_se force pass of one packing whose particles
// already live in device memory, for callers that cannot allocate (code under
// HIP graph capture, serving loops, allocation-free harnesses).  All working
// memory -- cell buckets, neighbour rows, per-workgroup partial sums, the
// packing's descriptors -- lives in one file-scope device array, `scratch`,
// of fixed size; the call allocates nothing and does not synchronise.  The
// force pass itself is the list kernel of sc_calc_se (k_list, k_reduce), so the
// results are those of that entry point on the list path with one lane per
// particle and 256-thread workgroups.
//
// The array's host-side name is only a stand-in for the device symbol: host
// code uses the address hipGetSymbolAddress returns for the current device,
// taken once per device on the first call there and kept.  Every call on a GPU
// shares that GPU's copy, so calls that overlap on one GPU race; each call
// clears what must start from zero (bucket counts, the error word, the rebuild
// flag) with a kernel of its own, so it is also safe to capture in a graph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "../../include/spherocyls.h"
#include "sc_internal.h"
#include "sc_physics.h"
#include "sc_status.h"

namespace sc {

template <int MODE, int LANES, bool TWO_POINT, bool MIXED, bool SE, bool BOX>
__global__ void k_list(ListArgs a);
__global__ void k_reduce(ReduceArgs a);
__global__ void k_bin_all(Bank b, const int* pack_of, const PackInfo* packs, const PackScal* scal,
                          int* cnt, int* slot, int* err, int ntot, int wrap_angles);

namespace {

constexpr int kNaBlock = 256;  // list-kernel workgroup, one lane per particle
constexpr int kNaMaxN = SC_NOALLOC_MAX_PARTICLES;
constexpr int kNaMaxCells = SC_NOALLOC_MAX_CELLS;
constexpr int kNaMaxBlocks = (kNaMaxN + kNaBlock - 1) / kNaBlock;

// Byte layout of the scratch array (each region 256-byte aligned).
struct Layout {
  size_t info, scal, pblocks, ranges, words, center, pack_sum, part, cnt, slot;
  size_t x, y, nx, ny, xb, yb, id, pack_of, nbr_n, nbr, total;
};

constexpr size_t up256(size_t v) { return (v + 255) / 256 * 256; }

constexpr Layout make_layout() {
  Layout l{};
  size_t o = 0;
  l.info = o; o = up256(o + sizeof(PackInfo));
  l.scal = o; o = up256(o + sizeof(PackScal));
  l.pblocks = o; o = up256(o + sizeof(int2) * kNaMaxBlocks);
  l.ranges = o; o = up256(o + sizeof(int2));
  l.words = o; o = up256(o + sizeof(int) * 4);  // [0] rebuild flag, [1] error bits
  l.center = o; o = up256(o + sizeof(double));
  l.pack_sum = o; o = up256(o + sizeof(double) * kNSum);
  l.part = o; o = up256(o + sizeof(double) * kNSum * kNaMaxBlocks);
  l.cnt = o; o = up256(o + sizeof(int) * kNaMaxCells);
  l.slot = o; o = up256(o + sizeof(int) * kNaMaxCells * kCapMin);
  l.x = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.y = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.nx = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.ny = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.xb = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.yb = o; o = up256(o + sizeof(double) * kNaMaxN);
  l.id = o; o = up256(o + sizeof(int) * kNaMaxN);
  l.pack_of = o; o = up256(o + sizeof(int) * kNaMaxN);
  l.nbr_n = o; o = up256(o + sizeof(int) * kNaMaxN);
  l.nbr = o; o = up256(o + sizeof(int) * (size_t)kNbrCap * kNaMaxN);
  l.total = o;
  return l;
}

constexpr Layout kLay = make_layout();
static_assert(kLay.total <= SC_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the stated size");

}  // namespace

// The scratch array (one copy per GPU, held until the process ends).
static __device__ __align__(256) unsigned char scratch[SC_NOALLOC_SCRATCH_BYTES];

namespace {

template <typename T>
__host__ __device__ T* at(unsigned char* base, size_t off) {
  return reinterpret_cast<T*>(base + off);
}

// Packing descriptors and the cell grid, from the particles on the device (one
// workgroup): configure_cells' rule (src/cpp/spherocyl_box.cpp:228-241) on
// Rmax + Amax; a grid of more than kNaMaxCells cells is coarsened (wider cells
// still hold every contact in the 3 x 3 block; only the bucket capacity can
// then run out, which the binning reports).  Clears the error word, raises the
// rebuild flag and writes the workgroup map of the list kernel.
__global__ void __launch_bounds__(256) k_na_setup(unsigned char* base, int n, double L,
                                                  double gamma, double pad, const double* A,
                                                  const double* R, int nblocks) {
  __shared__ double s_r[256], s_a[256];
  const int tid = threadIdx.x;
  double rm = 0.0, am = 0.0;
  for (int i = tid; i < n; i += blockDim.x) {
    rm = fmax(rm, R[i]);
    am = fmax(am, A[i]);
  }
  s_r[tid] = rm;
  s_a[tid] = am;
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if (tid < off) {
      s_r[tid] = fmax(s_r[tid], s_r[tid + off]);
      s_a[tid] = fmax(s_a[tid], s_a[tid + off]);
    }
    __syncthreads();
  }
  int2* pb = at<int2>(base, kLay.pblocks);
  for (int w = tid; w < nblocks; w += blockDim.x) pb[w] = make_int2(0, w * kNaBlock);
  if (tid != 0) return;
  const double reach = s_r[0] + s_a[0];
  const double wmin = 2.24 * reach + pad, hmin = 2 * reach + pad;
  int rows = max((int)(L / hmin), 1), cols = max((int)(L / wmin), 1);
  if (pad == 0) pad = fmax(L / cols - wmin, L / rows - hmin);
  while ((long long)rows * cols > kNaMaxCells) {
    rows = max(rows / 2, 1);
    cols = max(cols / 2, 1);
  }
  PackInfo P{};
  P.n = n;
  P.rows = rows;
  P.cols = cols;
  P.tile_cols = 1;
  P.cap = kCapMin;
  P.pad = pad;
  P.dt = 0.05;
  P.resize_scale = 1.0;
  P.inv_2n = 1.0 / (2.0 * n);
  *at<PackInfo>(base, kLay.info) = P;
  PackScal S{};
  S.Lx = S.Ly = L;
  S.gamma = gamma;
  S.prec = kFP64;
  *at<PackScal>(base, kLay.scal) = S;
  *at<int2>(base, kLay.ranges) = make_int2(0, nblocks);
  int* words = at<int>(base, kLay.words);
  words[0] = 1;  // rebuild the rows in the force pass
  words[1] = 0;  // error bits
}

// Copy of the caller's positions (binning wraps them), segment vectors, the
// identity ids and packing map, and zeroed bucket counts of the grid.
__global__ void k_na_prepare(unsigned char* base, int n, const double* x, const double* y,
                             const double* phi, const double* A) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const PackInfo& P = *at<PackInfo>(base, kLay.info);
  if (g < P.rows * P.cols) at<int>(base, kLay.cnt)[g] = 0;
  if (g >= n) return;
  at<double>(base, kLay.x)[g] = x[g];
  at<double>(base, kLay.y)[g] = y[g];
  double sn, cs;
  sincos(phi[g], &sn, &cs);
  at<double>(base, kLay.nx)[g] = A[g] * cs;
  at<double>(base, kLay.ny)[g] = A[g] * sn;
  at<int>(base, kLay.id)[g] = g;
  at<int>(base, kLay.pack_of)[g] = 0;
}

// Stress / energy record and residual to the caller, and the status word.
__global__ void k_na_finish(unsigned char* base, int n, double* se, int* status) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double* ps = at<double>(base, kLay.pack_sum);
  if (se) {
    for (int i = 0; i < kNSe; ++i) se[i] = ps[i];
    se[kNSe] = sqrt(ps[kSumR2] / (3.0 * n));
  }
  if (status) {
    const int e = at<int>(base, kLay.words)[1];
    *status = e == 0 ? SC_OK
                     : (e & kErrCellOverflow ? SC_ERR_CELL_OVERFLOW
                                             : (e & kErrList ? SC_ERR_LIST_OVERFLOW : SC_ERR_ARG));
  }
}

std::mutex g_mu;
unsigned char* g_ptr[64] = {};  // the scratch address on each device

// The scratch address on the current device, taken once per device.
int scratch_for_device(unsigned char** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return SC_ERR_NO_DEVICE;
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return SC_ERR_HIP;
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_ptr[dev];
  return SC_OK;
}

}  // namespace
}  // namespace sc

using namespace sc;

extern "C" int sc_calc_se_noalloc(int n, double L, double gamma, double pad, int contact_model,
                                  const double* x, const double* y, const double* phi,
                                  const double* A, const double* R, double* fx, double* fy,
                                  double* tau, double* se, int* status, void* stream) {
  if (n < 1 || n > SC_NOALLOC_MAX_PARTICLES || !(L > 0) || !(pad >= 0) ||
      !(std::fabs(gamma) <= 0.5) || !x || !y || !phi || !A || !R || !fx || !fy || !tau ||
      (contact_model != SC_CONTACT_CLOSEST && contact_model != SC_CONTACT_TWO_POINT))
    return SC_ERR_ARG;
  try {
    unsigned char* base = nullptr;
    int st = scratch_for_device(&base);
    if (st != SC_OK) return st;
    hipStream_t sm = static_cast<hipStream_t>(stream);
    const int nblocks = (n + kNaBlock - 1) / kNaBlock;
    hipLaunchKernelGGL(k_na_setup, dim3(1), dim3(256), 0, sm, base, n, L, gamma, pad, A, R,
                       nblocks);
    const int nprep = (std::max(n, kNaMaxCells) + 255) / 256;
    hipLaunchKernelGGL(k_na_prepare, dim3(nprep), dim3(256), 0, sm, base, n, x, y, phi, A);
    Bank b{};
    b.x = at<double>(base, kLay.x);
    b.y = at<double>(base, kLay.y);
    b.phi = const_cast<double*>(phi);
    b.R = const_cast<double*>(R);
    b.A = const_cast<double*>(A);
    b.nx = at<double>(base, kLay.nx);
    b.ny = at<double>(base, kLay.ny);
    b.id = at<int>(base, kLay.id);
    b.fx = fx;
    b.fy = fy;
    b.tau = tau;
    const PackInfo* packs = at<PackInfo>(base, kLay.info);
    const PackScal* scal = at<PackScal>(base, kLay.scal);
    int* words = at<int>(base, kLay.words);
    int* cnt = at<int>(base, kLay.cnt);
    int* slot = at<int>(base, kLay.slot);
    hipLaunchKernelGGL(k_bin_all, dim3((n + 255) / 256), dim3(256), 0, sm, b,
                       at<int>(base, kLay.pack_of), packs, scal, cnt, slot, words + 1, n, 0);
    ListArgs a{};
    a.packs = packs;
    a.scal = scal;
    a.scal_next = nullptr;
    a.pblocks = at<int2>(base, kLay.pblocks);
    a.b = b;
    a.ex = b.x;
    a.ey = b.y;
    a.enx = b.nx;
    a.eny = b.ny;
    a.cnt = cnt;
    a.slot = slot;
    a.ncells_total = 0;
    a.nbr = at<int>(base, kLay.nbr);
    a.nbr_n = at<int>(base, kLay.nbr_n);
    a.xb = at<double>(base, kLay.xb);
    a.yb = at<double>(base, kLay.yb);
    a.list_center = at<double>(base, kLay.center);
    a.flag_in = words;
    a.gamma_window = 0.05;
    a.part = at<double>(base, kLay.part);
    a.err = words + 1;
    a.ntot = n;
    a.npack = 1;
    a.se_always = 1;
    a.do_resid = 1;
    const bool two = contact_model == SC_CONTACT_TWO_POINT;
    auto fn = two ? &k_list<kModeForces, 1, true, false, true, false>
                  : &k_list<kModeForces, 1, false, false, true, false>;
    hipLaunchKernelGGL(fn, dim3(nblocks), dim3(kNaBlock), sizeof(int) * kHitCap * kNaBlock, sm,
                       a);
    ReduceArgs r{};
    r.packs = packs;
    r.scal = scal;
    r.part = a.part;
    r.ranges = at<int2>(base, kLay.ranges);
    r.pack_sum = at<double>(base, kLay.pack_sum);
    r.se_always = 1;
    r.npack = 1;
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(256), 0, sm, r);
    hipLaunchKernelGGL(k_na_finish, dim3(1), dim3(64), 0, sm, base, n, se, status);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? SC_OK : SC_ERR_HIP;
  } catch (...) {
    return current_exception_status(nullptr);
  }
}
