// nufftx_spread_noalloc: a spreading entry point that allocates no device memory.
//
// It runs the same bin-sort and spreading kernels as nufftx_makeplan_grid + nufftx_setpts +
// nufftx_spread (one transform), with every working buffer of the sort and the spread (keys,
// histogram, segment sums, permutation, inverse permutation, subproblem list, packed point
// records, strength bounds) carved out of one file-scope device array.  The array's device
// address is taken once per GPU with hipGetSymbolAddress (each GPU holds its own copy, at its
// own address); host code never passes the array's name to a kernel.  Nothing here calls
// hipMalloc, hipFree or an allocating library routine: the host-side bookkeeping (a PointSet
// whose buffers point into the array, with their capacities set) makes every grow_buffer a
// no-op.  Calls that overlap on one GPU share the array and race; the segment sums, which
// the sort expects at zero, are cleared by a fill kernel at every call.
#include <mutex>

#include "../../include/nufftx.h"
#include "internal.hpp"

namespace nufftx {

namespace {

constexpr size_t kScratchBytes = (size_t)NUFFTX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// device address of this GPU's scratch array (the current device's)
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return fail(NUFFTX_ERR_HIP, "nufftx_spread_noalloc: no current device");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(NUFFTX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return NUFFTX_OK;
}

// bump allocator over the scratch array (256-byte aligned pieces)
struct Carve {
  unsigned char* base;
  size_t used = 0;
  template <typename T>
  T* take(int64_t n, int64_t& cap) {
    used = (used + 255) / 256 * 256;
    T* p = reinterpret_cast<T*>(base + used);
    used += (size_t)std::max<int64_t>(n, 1) * sizeof(T);
    cap = std::max<int64_t>(n, 1);
    return p;
  }
};

}  // namespace

}  // namespace nufftx

using namespace nufftx;

extern "C" int nufftx_spread_noalloc(int dim, const int64_t* nf, int nspread, double beta, double kc, int precision,
                                     int64_t M, const void* x, const void* y, const void* z, const void* c, void* fw,
                                     void* stream) {
  if (dim < 1 || dim > 3) return fail(NUFFTX_ERR_INVALID_ARG, "dim must be 1, 2 or 3");
  if (precision != NUFFTX_F32 && precision != NUFFTX_F64) return fail(NUFFTX_ERR_INVALID_ARG, "bad precision");
  if (!nf) return fail(NUFFTX_ERR_INVALID_ARG, "nf is null");
  if (nspread < 2 || nspread > 16 || (nspread & 1)) return fail(NUFFTX_ERR_INVALID_ARG, "nspread must be even in [2, 16]");
  if (M < 0 || M > NUFFTX_NOALLOC_MAX_POINTS)
    return fail(NUFFTX_ERR_INVALID_ARG, "nufftx_spread_noalloc: M above NUFFTX_NOALLOC_MAX_POINTS");
  for (int d = 0; d < dim; ++d)
    if (nf[d] < 1 || nf[d] > (1LL << 30)) return fail(NUFFTX_ERR_INVALID_ARG, "nf out of range");
  if (!fw) return fail(NUFFTX_ERR_INVALID_ARG, "fw is null");
  const void* xs[3] = {x, y, z};
  for (int d = 0; d < dim; ++d)
    if (M > 0 && !xs[d]) return fail(NUFFTX_ERR_INVALID_ARG, "missing coordinate array");
  if (M > 0 && !c) return fail(NUFFTX_ERR_INVALID_ARG, "c is null");
  hipStream_t s = static_cast<hipStream_t>(stream);

  Geometry g;
  g.dim = dim;
  g.ns = nspread;
  for (int d = 0; d < 3; ++d) g.nf[d] = d < dim ? (int)nf[d] : 1;
  choose_bins(g, precision, lds_budget_bytes());
  SortTuning tune;
  const bool slot = win2d_geometry(g, precision, tune.maxsub);

  void* sp = nullptr;
  if (int r = scratch_address(&sp)) return r;
  PointSet ps;
  ps.M = M;
  for (int d = 0; d < 3; ++d) ps.coord[d] = d < dim ? xs[d] : nullptr;
  ps.coord_f64 = precision == NUFFTX_F64;
  const SortSizes zs = sort_sizes(M, g, tune, 0);
  Carve cv{static_cast<unsigned char*>(sp)};
  int64_t cap_nsub = 0;
  ps.perm = cv.take<int32_t>(M, ps.cap_M);
  ps.keys = cv.take<int32_t>(M, ps.cap_keys);
  ps.hist = cv.take<int32_t>(zs.nhist, ps.cap_hist);
  ps.scan_part = cv.take<int32_t>(zs.nparts, ps.cap_part);
  ps.scan_sum = cv.take<int32_t>(zs.nparts, ps.cap_sum);
  ps.bin_start = cv.take<int32_t>((int64_t)g.nbins + 1, ps.cap_bins);
  ps.sub_start = cv.take<int32_t>((int64_t)g.nbins + 1, ps.cap_sub_start);
  ps.subprob = cv.take<int4>(zs.max_sub, ps.cap_sub);
  ps.nsub = cv.take<int32_t>(1, cap_nsub);
  // strength bound of the LDS-atomic kernels: one value and its per-block partial maxima
  ps.cmax = cv.take<float>(1 + 1024, ps.cap_cmax);
  if (slot) {
    ps.slot = cv.take<int32_t>(M, ps.cap_slot);
    ps.packed = cv.take<uint4>(M, ps.cap_packed);
  }
  if (cv.used > kScratchBytes)
    return fail(NUFFTX_ERR_INVALID_ARG, "nufftx_spread_noalloc: the problem needs " + std::to_string(cv.used) +
                                            " bytes of working memory, above NUFFTX_NOALLOC_SCRATCH_BYTES");
  // the sort expects its segment sums at zero (a fill kernel: no memset node under capture)
  if (launch_zero(ps.scan_sum, ps.cap_sum * (int64_t)sizeof(int32_t), s) != hipSuccess)
    return fail(NUFFTX_ERR_HIP, std::string("nufftx_spread_noalloc: ") + hipGetErrorString(hipGetLastError()));
  std::string err;
  hipError_t e = pointset_sort(ps, g, tune, s, err, 0, 0, slot, 1);
  if (e != hipSuccess) {
    if (!err.empty()) return fail(NUFFTX_ERR_INVALID_ARG, err);
    return fail(NUFFTX_ERR_HIP, std::string("nufftx_spread_noalloc bin-sort: ") + hipGetErrorString(e));
  }
  if (!ps.retired.empty()) return fail(NUFFTX_ERR_STATE, "nufftx_spread_noalloc: a buffer outgrew its scratch slice");
  GridOpArgs ga;
  ga.zero_out = true;  // fw = spread(c)
  ga.ps = &ps;
  ga.g = &g;
  ga.prec = precision;
  ga.n_trans = 1;
  ga.beta = beta;
  ga.c = kc;
  if ((e = launch_spread(ga, c, fw, s)) != hipSuccess)
    return fail(NUFFTX_ERR_HIP, std::string("nufftx_spread_noalloc spread: ") + hipGetErrorString(e));
  if (!ps.retired.empty()) return fail(NUFFTX_ERR_STATE, "nufftx_spread_noalloc: a buffer outgrew its scratch slice");
  return NUFFTX_OK;
}
