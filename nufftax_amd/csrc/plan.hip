// Plans, host-side FP64 parameter setup, rocFFT and the C ABI (include/nufftx.h).
#include <rocfft/rocfft.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nufftx.h"
#include "internal.hpp"

namespace nufftx {

// ------------------------------------------------------------------ errors
namespace {
thread_local std::string g_last_error;
}  // namespace

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

// ------------------------------------------------------------------ device allocations
namespace {
std::atomic<int64_t> g_dev_allocs{0};
}  // namespace

hipError_t dev_malloc(void** p, size_t bytes) {
  g_dev_allocs.fetch_add(1, std::memory_order_relaxed);
  return hipMalloc(p, bytes);
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail(NUFFTX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ host math
// nufftax/core/kernel.py:165-206
void kernel_params(double tol, double upsampfac, int& ns, double& beta, double& c) {
  const double log_tol = -std::log10(std::max(tol, 1e-16));
  ns = (int)std::ceil(log_tol + 1);
  ns = std::max(2, std::min(ns, 16));
  if (ns % 2 == 1) ns += 1;
  const double sigma = upsampfac;
  const double inner = (double)(ns * ns) / (sigma * sigma) * (sigma - 0.5) * (sigma - 0.5) - 0.8;
  beta = inner > 0 ? M_PI * std::sqrt(inner) : 2.3 * ns;
  c = 4.0 / (double)(ns * ns);
}

// The polynomial ES kernel (common.hpp): P_k interpolates phi(z0_k + v/2), z0_k = 1/2 + k - ns/2,
// at the 10 Chebyshev points of [-1, 1] (FP64, degree 9), converted to monomials of v and split
// into even and odd parts, rounded to FP32; then every interval (P_k(v) for k < ns/2,
// P_k(-v) for the mirrored ones) is evaluated as the kernels do (FP32 FMAs) at 257 points and
// compared with the exact phi (nufftax/core/kernel.py:79-111).
namespace {
EsPoly es_poly_fit(int ns, double beta, double c) {
  EsPoly p{};
  p.on = 0;
  if (ns != kEsPolyNs) return p;
  constexpr int N = 10;  // degree 9
  auto phi = [&](double z) {
    const double a = 1.0 - c * z * z;
    return a >= 0.0 ? std::exp(beta * (std::sqrt(a) - 1.0)) : 0.0;
  };
  double tm[N][N] = {};  // tm[n][d]: coefficient of v^d in the Chebyshev polynomial T_n
  tm[0][0] = 1.0;
  tm[1][1] = 1.0;
  for (int n = 2; n < N; ++n)
    for (int d = 0; d < N; ++d) tm[n][d] = (d > 0 ? 2.0 * tm[n - 1][d - 1] : 0.0) - tm[n - 2][d];
  for (int k = 0; k < ns / 2; ++k) {
    const double z0 = 0.5 + k - 0.5 * ns;
    double f[N], a[N];
    for (int j = 0; j < N; ++j) f[j] = phi(z0 + 0.5 * std::cos(M_PI * (j + 0.5) / N));
    for (int n = 0; n < N; ++n) {
      double sum = 0.0;
      for (int j = 0; j < N; ++j) sum += f[j] * std::cos(n * M_PI * (j + 0.5) / N);
      a[n] = sum * (n == 0 ? 1.0 : 2.0) / N;
    }
    for (int d = 0; d < N; ++d) {
      double m = 0.0;
      for (int n = d; n < N; ++n) m += a[n] * tm[n][d];
      (d % 2 == 0 ? p.e[k][d / 2] : p.o[k][d / 2]) = (float)m;
    }
  }
  double maxerr = 0.0;
  for (int q = 0; q <= 256; ++q) {
    const float v = -1.0f + (float)q / 128.0f, u = v * v;
    for (int k = 0; k < ns / 2; ++k) {
      float ev = p.e[k][4], od = p.o[k][4];
      for (int d = 3; d >= 0; --d) {
        ev = std::fma(ev, u, p.e[k][d]);
        od = std::fma(od, u, p.o[k][d]);
      }
      const double z0 = 0.5 + k - 0.5 * ns, z1 = 0.5 + (ns - 1 - k) - 0.5 * ns;
      maxerr = std::max(maxerr, std::fabs((double)std::fma(v, od, ev) - phi(z0 + 0.5 * (double)v)));
      maxerr = std::max(maxerr, std::fabs((double)std::fma(-v, od, ev) - phi(z1 + 0.5 * (double)v)));
    }
  }
  p.on = maxerr <= kEsPolyTol ? 1 : 0;
  return p;
}
}  // namespace

EsPoly es_poly_table(int ns, double beta, double c) {
  struct Entry {
    int ns;
    double beta, c;
    EsPoly p;
  };
  static std::mutex mu;
  static std::vector<Entry> fits;  // one per parameter set the process has used (a few)
  EsPoly p{};
  p.on = 0;
  if (ns != kEsPolyNs) return p;
  if (const char* e = std::getenv("NUFFTX_ES_POLY"))
    if (e[0] == '0') return p;
  std::lock_guard<std::mutex> lk(mu);
  for (const Entry& e : fits)
    if (e.ns == ns && e.beta == beta && e.c == c) return e.p;
  p = es_poly_fit(ns, beta, c);
  if (fits.size() < 256) fits.push_back(Entry{ns, beta, c, p});
  return p;
}

// nufftax/utils/grid.py:6-33
int64_t next_smooth_int(int64_t n) {
  if (n <= 1) return 1;
  for (;; ++n) {
    int64_t v = n;
    for (int p : {2, 3, 5})
      while (v % p == 0) v /= p;
    if (v == 1) return n;
  }
}

// nufftax/utils/grid.py:36-58
int64_t grid_size(int64_t n_modes, double upsampfac, int ns) {
  return next_smooth_int(std::max<int64_t>((int64_t)std::ceil(upsampfac * (double)n_modes), 2 * ns));
}

// Trapezoid rule of nufftax/core/kernel.py:237-256 (also nufft3.py:195-206), FP64.
QuadRule quad_rule(int ns, double beta, double c) {
  QuadRule q;
  const double j2 = ns / 2.0;
  q.n = std::max((int)(4 + 3.0 * j2), 20);
  const double dz = (2.0 * j2) / (q.n - 1);
  for (int i = 0; i < q.n; ++i) {
    const double z = -j2 + dz * i;
    q.z[i] = z;
    const double arg = 1.0 - c * z * z;
    const double phi = arg >= 0 ? std::exp(beta * (std::sqrt(std::max(arg, 0.0)) - 1.0)) : 0.0;
    const double w = (i == 0 || i == q.n - 1) ? dz / 2 : dz;
    q.wphi[i] = w * phi;
  }
  for (int i = q.n; i < 40; ++i) q.z[i] = q.wphi[i] = 0.0;
  return q;
}

// phihat[k], k = 0..nf/2 (kernel.py:209-260)
std::vector<double> phihat(int64_t nf, const QuadRule& q) {
  std::vector<double> out(nf / 2 + 1);
  for (int64_t k = 0; k <= nf / 2; ++k) {
    double s = 0;
    for (int i = 0; i < q.n; ++i) s += q.wphi[i] * std::cos(2.0 * M_PI * (double)k * q.z[i] / (double)nf);
    out[k] = s;
  }
  return out;
}

// (-1)^k / phihat[|k|] in output mode order (deconvolve.py:27-56)
std::vector<double> deconv_factors(const std::vector<double>& ph, int64_t n, int modeord) {
  const int64_t kmin = -(n / 2), kmax = (n - 1) / 2;
  std::vector<double> f(n);
  for (int64_t m = 0; m < n; ++m) {
    int64_t k = modeord == 0 ? m + kmin : (m <= kmax ? m : m - n);
    const double sgn = (k % 2 == 0) ? 1.0 : -1.0;
    f[m] = sgn / ph[(size_t)std::llabs(k)];
  }
  return f;
}

// Sort/subproblem tuning, overridable for sweeps (NUFFTX_MAXSUB, NUFFTX_SORT_CHUNK).
SortTuning default_tuning() {
  SortTuning t;
  if (const char* e = std::getenv("NUFFTX_MAXSUB")) t.maxsub = std::max(64, std::atoi(e));
  if (const char* e = std::getenv("NUFFTX_SORT_CHUNK")) t.chunk = std::max(256, std::atoi(e));
  return t;
}

int lds_budget_bytes() {
  static int budget = [] {
    int b = 65536;
    if (const char* e = std::getenv("NUFFTX_LDS_BUDGET")) b = std::max(4096, std::atoi(e));
    int dev = 0, mx = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&mx, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && mx > 0)
      b = std::min(b, mx);
    return b;
  }();
  return budget;
}

// Largest dynamic LDS allocation of one block on the current device.
int lds_device_max_bytes() {
  static int mx = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      return v;
    return 65536;
  }();
  return mx;
}

// Bin sizes per dimension: tiles (bin + ns halo) must fit the LDS budget and the
// bin count the LDS histogram (16384 counters).
void choose_bins(Geometry& g, int prec, int budget) {
  const size_t cell = 16;  // spread tiles hold two 8-byte accumulators per cell
  int bs[3];
  if (g.dim == 1) {
    bs[0] = 1024;
    bs[1] = bs[2] = 1;
  } else if (g.dim == 2) {
    // FP32, ns <= 8: bins of 4 x 4 sub-bins of 16 - ns cells for the MFMA spread
    const int b2 = (prec == kF32 && g.ns <= 8) ? 4 * (16 - g.ns) : 32;
    bs[0] = b2;
    bs[1] = b2;
    bs[2] = 1;
  } else {
    bs[0] = 16;
    bs[1] = 16;
    bs[2] = 8;
  }
  // 3D FP32, ns <= 8: bins of 2 x 1 x 2 sub-bins of the 3D window spread (8 x 8 x 4 support
  // starts), kept at that size whatever the budget: the window kernel's exact tiles (two
  // 64-bit words per cell) fit the LDS at this size, and the bank kernel's tiles still fit
  // the device's LDS.  On 8M clustered points (128^3 modes) the interpolation and the bank
  // spread (ns < 8) run as fast with these bins as with 16 x 16 x 8 (type 2 at eps 1e-4:
  // 1.29 against 1.31 ms).
  const bool win3 = g.dim == 3 && prec == kF32 && g.ns <= 8 && g.nf[0] >= 16 && g.nf[1] >= 8 && g.nf[2] >= 8;
  if (win3) {
    bs[0] = 16;
    bs[1] = 8;
    bs[2] = 8;
  }
  for (int d = 0; d < 3; ++d) bs[d] = d < g.dim ? std::max(1, std::min(bs[d], g.nf[d])) : 1;
  auto tile_bytes = [&](const int* b) {
    size_t t = 1;
    for (int d = 0; d < g.dim; ++d) t *= (size_t)(b[d] + g.ns);
    return t * cell;
  };
  auto nbins = [&](const int* b) {
    int64_t n = 1;
    for (int d = 0; d < g.dim; ++d) n *= (g.nf[d] + b[d] - 1) / b[d];
    return n;
  };
  while (!win3 && tile_bytes(bs) > (size_t)budget) {
    int dmax = 0;
    for (int d = 1; d < g.dim; ++d)
      if (bs[d] > bs[dmax]) dmax = d;
    if (bs[dmax] <= 4) break;
    bs[dmax] /= 2;
  }
  while (nbins(bs) > 16384) {
    int dmin = -1;
    for (int d = 0; d < g.dim; ++d)
      if (bs[d] < g.nf[d] && (dmin < 0 || bs[d] < bs[dmin])) dmin = d;
    if (dmin < 0) break;
    bs[dmin] = std::min(g.nf[dmin], bs[dmin] * 2);
  }
  // (doubling keeps the window bins whole numbers of sub-bins; the window kernel checks the
  // rest: at most 16 sub-bins, tiles under 128 cells a side)
  g.nbins = 1;
  g.tile_cells = 1;
  for (int d = 0; d < 3; ++d) {
    g.bs[d] = bs[d];
    g.nb[d] = d < g.dim ? (g.nf[d] + bs[d] - 1) / bs[d] : 1;
    g.tile[d] = d < g.dim ? bs[d] + g.ns : 1;
    g.nbins *= g.nb[d];
    g.tile_cells *= g.tile[d];
  }
}

// ------------------------------------------------------------------ rocFFT
namespace {

std::once_flag g_fft_once;
rocfft_status g_fft_setup = rocfft_status_success;

struct FftPlan {
  rocfft_plan fwd = nullptr, bwd = nullptr;
  rocfft_execution_info info = nullptr;
  void* work = nullptr;
};

void fft_free(FftPlan& p) {
  if (p.fwd) rocfft_plan_destroy(p.fwd);
  if (p.bwd) rocfft_plan_destroy(p.bwd);
  if (p.info) rocfft_execution_info_destroy(p.info);
  if (p.work) (void)hipFree(p.work);
  p = FftPlan();
}

int fft_make(FftPlan& p, int dim, const int nf[3], int64_t n_trans, int prec) {
  std::call_once(g_fft_once, [] { g_fft_setup = rocfft_setup(); });
  if (g_fft_setup != rocfft_status_success) return fail(NUFFTX_ERR_FFT, "rocfft_setup failed");
  size_t lengths[3] = {(size_t)nf[0], (size_t)nf[1], (size_t)nf[2]};
  const rocfft_precision rp = prec == kF32 ? rocfft_precision_single : rocfft_precision_double;
  rocfft_status st = rocfft_plan_create(&p.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward, rp,
                                        (size_t)dim, lengths, (size_t)n_trans, nullptr);
  if (st != rocfft_status_success) return fail(NUFFTX_ERR_FFT, "rocfft_plan_create (forward) failed");
  st = rocfft_plan_create(&p.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, rp, (size_t)dim,
                          lengths, (size_t)n_trans, nullptr);
  if (st != rocfft_status_success) return fail(NUFFTX_ERR_FFT, "rocfft_plan_create (inverse) failed");
  size_t w1 = 0, w2 = 0;
  rocfft_plan_get_work_buffer_size(p.fwd, &w1);
  rocfft_plan_get_work_buffer_size(p.bwd, &w2);
  if (rocfft_execution_info_create(&p.info) != rocfft_status_success)
    return fail(NUFFTX_ERR_FFT, "rocfft_execution_info_create failed");
  const size_t w = std::max(w1, w2);
  if (w > 0) {
    HIPCHK(dev_malloc(&p.work, w));
    if (rocfft_execution_info_set_work_buffer(p.info, p.work, w) != rocfft_status_success)
      return fail(NUFFTX_ERR_FFT, "rocfft_execution_info_set_work_buffer failed");
  }
  return NUFFTX_OK;
}

// isign > 0: unnormalised inverse (ifft * N, nufft1.py:81); isign < 0: forward.
int fft_exec(FftPlan& p, void* data, int isign, hipStream_t s) {
  if (rocfft_execution_info_set_stream(p.info, s) != rocfft_status_success)
    return fail(NUFFTX_ERR_FFT, "rocfft_execution_info_set_stream failed");
  void* in[1] = {data};
  if (rocfft_execute(isign > 0 ? p.bwd : p.fwd, in, nullptr, p.info) != rocfft_status_success)
    return fail(NUFFTX_ERR_FFT, "rocfft_execute failed");
  return NUFFTX_OK;
}

template <typename T>
int dev_alloc(T*& p, size_t n) {
  if (p) return NUFFTX_OK;
  if (dev_malloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    p = nullptr;
    return fail(NUFFTX_ERR_ALLOC, "hipMalloc failed (" + std::to_string(n * sizeof(T)) + " bytes)");
  }
  return NUFFTX_OK;
}

// Grow a plan-owned buffer without freeing the outgrown one (see grow_buffer).
template <typename T>
int dev_grow(T*& p, int64_t& cap, int64_t n, std::vector<void*>& retired) {
  if (p && n <= cap) return NUFFTX_OK;
  if (grow_buffer(p, cap, n, retired) != hipSuccess)
    return fail(NUFFTX_ERR_ALLOC, "hipMalloc failed (" + std::to_string(n * sizeof(T)) + " bytes)");
  return NUFFTX_OK;
}

int upload_factors(void*& dst, const std::vector<double>& f, int prec) {
  const size_t n = f.size();
  if (prec == kF32) {
    std::vector<float> h(f.begin(), f.end());
    float* d = nullptr;
    if (int r = dev_alloc(d, n)) return r;
    HIPCHK(hipMemcpy(d, h.data(), n * sizeof(float), hipMemcpyHostToDevice));
    dst = d;
  } else {
    double* d = nullptr;
    if (int r = dev_alloc(d, n)) return r;
    HIPCHK(hipMemcpy(d, f.data(), n * sizeof(double), hipMemcpyHostToDevice));
    dst = d;
  }
  return NUFFTX_OK;
}

}  // namespace

}  // namespace nufftx

using namespace nufftx;

struct nufftx_plan_s {
  int type = 1;  // 0 = grid operator plan
  int dim = 1;
  int prec = kF32;
  int isign = 1;
  int modeord = 0;
  int64_t n_trans = 1;
  double eps = 1e-6, upsampfac = 2.0;
  int ns = 8;
  double beta = 0, kc = 0;
  int64_t n_modes[3] = {1, 1, 1};
  Geometry g, gi;
  hipStream_t stream = nullptr;
  PointSet ps, pt;
  SortTuning tune;
  bool pts_set = false;
  bool cached = false;  // owned by the plan cache (nufftx_plan_acquire): released, never destroyed
  void* grid = nullptr;
  void* grid_in = nullptr;
  void* fac[3] = {nullptr, nullptr, nullptr};
  void* fac_in[3] = {nullptr, nullptr, nullptr};
  FftPlan fft, fft_in;
  Fft2Plan fft2;          // fused FFT + deconvolution (2D FP32, power-of-two grids)
  bool fft2_pad = false;      // modes -> grid (type 2 input, adjoint of type 1) fused
  bool fft2_shuffle = false;  // grid -> modes (type 1 output) fused
  QuadRule quad;
  // type 3
  T3State* t3 = nullptr;
  double* red = nullptr;
  double* xr[3] = {nullptr, nullptr, nullptr};
  double* sr[3] = {nullptr, nullptr, nullptr};
  double* pre = nullptr;
  double* post = nullptr;
  void* cbuf = nullptr;
  int64_t cap_xr[3] = {0, 0, 0}, cap_sr[3] = {0, 0, 0}, cap_pre = 0, cap_post = 0, cap_cbuf = 0;
  std::vector<void*> retired;  // outgrown type-3 buffers, freed with the plan
  int64_t M = 0, N = 0;
  size_t csize() const { return prec == kF32 ? 8 : 16; }
};

namespace {

void plan_free(nufftx_plan p) {
  pointset_free(p->ps);
  pointset_free(p->pt);
  fft_free(p->fft);
  fft2_plan_free(p->fft2);
  fft_free(p->fft_in);
  void* bufs[] = {p->grid, p->grid_in, p->fac[0], p->fac[1], p->fac[2], p->fac_in[0], p->fac_in[1], p->fac_in[2],
                  p->t3, p->red, p->xr[0], p->xr[1], p->xr[2], p->sr[0], p->sr[1], p->sr[2], p->pre, p->post,
                  p->cbuf};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  for (void* b : p->retired) (void)hipFree(b);
  p->retired.clear();
}

// nufftx_makeplan's argument checks (no device call)
int check_plan_args(int type, int dim, const int64_t* n_modes, int64_t n_trans, double eps, double upsampfac,
                    int precision, int modeord) {
  if (type < 1 || type > 3) return fail(NUFFTX_ERR_INVALID_ARG, "type must be 1, 2 or 3");
  if (dim < 1 || dim > 3) return fail(NUFFTX_ERR_INVALID_ARG, "dim must be 1, 2 or 3");
  if (!n_modes) return fail(NUFFTX_ERR_INVALID_ARG, "n_modes is null");
  if (n_trans < 1 || n_trans > 65535) return fail(NUFFTX_ERR_INVALID_ARG, "n_trans must be in [1, 65535]");
  if (!(eps > 0)) return fail(NUFFTX_ERR_INVALID_ARG, "eps must be positive");
  if (!(upsampfac > 1.0) || upsampfac > 16.0) return fail(NUFFTX_ERR_INVALID_ARG, "upsampfac must be in (1, 16]");
  if (precision != NUFFTX_F32 && precision != NUFFTX_F64) return fail(NUFFTX_ERR_INVALID_ARG, "bad precision");
  if (modeord != 0 && modeord != 1) return fail(NUFFTX_ERR_INVALID_ARG, "modeord must be 0 or 1");
  for (int d = 0; d < dim; ++d)
    if (n_modes[d] < 1 || n_modes[d] > (1LL << 28)) return fail(NUFFTX_ERR_INVALID_ARG, "n_modes out of range");
  return NUFFTX_OK;
}

// nufftx_makeplan_grid's argument checks (no device call)
int check_grid_args(int dim, const int64_t* nf, int nspread, int64_t n_trans, int precision) {
  if (dim < 1 || dim > 3 || !nf) return fail(NUFFTX_ERR_INVALID_ARG, "dim must be 1, 2 or 3");
  if (nspread < 2 || nspread > 16 || nspread % 2) return fail(NUFFTX_ERR_INVALID_ARG, "nspread must be even in [2,16]");
  if (n_trans < 1 || n_trans > 65535) return fail(NUFFTX_ERR_INVALID_ARG, "n_trans must be in [1, 65535]");
  if (precision != NUFFTX_F32 && precision != NUFFTX_F64) return fail(NUFFTX_ERR_INVALID_ARG, "bad precision");
  for (int d = 0; d < dim; ++d)
    if (nf[d] < 1 || nf[d] > (1LL << 30)) return fail(NUFFTX_ERR_INVALID_ARG, "nf out of range");
  return NUFFTX_OK;
}

int check_plan(nufftx_plan p) {
  if (!p) return fail(NUFFTX_ERR_INVALID_ARG, "null plan");
  return NUFFTX_OK;
}

// interp: the point set is interpolated to (type 2, grid operators, type-3 targets).
int sort_points(nufftx_plan p, PointSet& ps, const Geometry& g, bool interp, bool spread) {
  std::string err;
  SortTuning tune = p->tune;
  // point sets that the window kernels spread or interpolate get subproblems of up to 4096
  // points (one block of the window kernels per bin of the 2D radial workload)
  if (((spread && win_spread_enabled()) || (interp && win_interp_enabled())) && win2d_geometry(g, p->prec, kWinMaxSub) &&
      !std::getenv("NUFFTX_MAXSUB"))
    tune.maxsub = kWinMaxSub;
  // the 3D window spread takes subproblems of up to 2048 points
  if (spread && win_spread_enabled() && win3d_geometry(g, p->prec, kWin3MaxSub) && !std::getenv("NUFFTX_MAXSUB"))
    tune.maxsub = kWin3MaxSub;
  const bool subsort = subsort_applicable(tune, g, p->prec);
  // records pay for their cost in the sort only where the point set is interpolated;
  // the matrix-core kernels need the full records with sub-bins
  const int rec_kind = g.dim == 2 && p->prec == kF32 ? (subsort ? 1 : interp ? 2 : 0) : 0;
  // point sets that the matrix-core window kernel spreads get the inverse permutation (and,
  // when spreading is their forward use, the packed-record buffer)
  const bool slot = win2d_geometry(g, p->prec, tune.maxsub) ||
                    (spread && win_spread_enabled() && win3d_geometry(g, p->prec, tune.maxsub));
  hipError_t e = pointset_sort(ps, g, tune, p->stream, err, rec_kind, subsort ? 16 - g.ns : 0, slot,
                               spread ? p->n_trans : 0);
  if (e != hipSuccess) {
    if (!err.empty()) return fail(NUFFTX_ERR_INVALID_ARG, err);
    return fail(NUFFTX_ERR_HIP, std::string("bin-sort: ") + hipGetErrorString(e));
  }
  if (subsort && ps.has_recs) {
    e = pointset_subsort(ps, g, p->stream);
    if (e != hipSuccess) return fail(NUFFTX_ERR_HIP, std::string("sub-bin sort: ") + hipGetErrorString(e));
  }
  // the FP32 spreading kernels' strength bound (n_trans values and the per-block partial
  // maxima) is sized here, so that execute allocates nothing for a point set of this size
  if (p->prec == kF32) {
    const int64_t nb = std::min<int64_t>((ps.M + 1023) / 1024, kAbsmaxBlocks);
    if (grow_buffer(ps.cmax, ps.cap_cmax, p->n_trans * (1 + std::max<int64_t>(nb, 1)), ps.retired) != hipSuccess)
      return fail(NUFFTX_ERR_ALLOC, "strength bound buffer");
  }
  return NUFFTX_OK;
}

GridOpArgs gop(nufftx_plan p, const PointSet& ps, const Geometry& g) {
  GridOpArgs a;
  a.ps = &ps;
  a.g = &g;
  a.prec = p->prec;
  a.n_trans = p->n_trans;
  a.beta = p->beta;
  a.c = p->kc;
  a.tol = p->eps;
  return a;
}

DeconvArgs dargs(nufftx_plan p, bool inner) {
  DeconvArgs a;
  a.dim = p->dim;
  a.prec = p->prec;
  a.n_trans = p->n_trans;
  a.modeord = inner ? 0 : p->modeord;
  for (int d = 0; d < 3; ++d) {
    a.nf[d] = inner ? p->gi.nf[d] : p->g.nf[d];
    a.nm[d] = inner ? p->g.nf[d] : (int)p->n_modes[d];
    a.fac[d] = inner ? p->fac_in[d] : p->fac[d];
  }
  return a;
}

#define LAUNCH(expr)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(NUFFTX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// fine grid -> FFT -> deconvolved modes f (type 1 output stage)
int grid_to_modes(nufftx_plan p, void* f, int isign) {
  if (p->fft2_shuffle) {
    LAUNCH(fft2_shuffle(p->fft2, dargs(p, false), p->grid, f, isign, p->stream));
    return NUFFTX_OK;
  }
  if (int r = fft_exec(p->fft, p->grid, isign, p->stream)) return r;
  LAUNCH(launch_deconv_shuffle(dargs(p, false), p->grid, f, p->stream));
  return NUFFTX_OK;
}

// modes f -> deconvolved, padded -> FFT -> fine grid (type 2 input stage)
int modes_to_grid(nufftx_plan p, const void* f, int isign) {
  if (p->fft2_pad) {
    LAUNCH(fft2_pad(p->fft2, dargs(p, false), f, p->grid, isign, p->stream));
    return NUFFTX_OK;
  }
  LAUNCH(launch_deconv_pad(dargs(p, false), f, p->grid, p->stream));
  return fft_exec(p->fft, p->grid, isign, p->stream);
}

int exec_type1(nufftx_plan p, const void* c, void* f, int isign) {
  GridOpArgs ga = gop(p, p->ps, p->g);
  ga.zero_out = true;
  LAUNCH(launch_spread(ga, c, p->grid, p->stream));
  return grid_to_modes(p, f, isign);
}

int exec_type2(nufftx_plan p, void* c, const void* f, int isign) {
  if (int r = modes_to_grid(p, f, isign)) return r;
  LAUNCH(launch_interp(gop(p, p->ps, p->g), p->grid, c, p->stream));
  return NUFFTX_OK;
}

int exec_type3(nufftx_plan p, const void* c, void* f) {
  LAUNCH(launch_cmul_pointwise(p->prec, p->n_trans, p->M, c, p->pre, p->cbuf, p->stream));
  GridOpArgs ga = gop(p, p->ps, p->g);
  ga.zero_out = true;
  LAUNCH(launch_spread(ga, p->cbuf, p->grid, p->stream));
  // inner type 2 on the spread grid as CMCL modes (nufft3.py:292, 387, 494-502)
  LAUNCH(launch_deconv_pad(dargs(p, true), p->grid, p->grid_in, p->stream));
  if (int r = fft_exec(p->fft_in, p->grid_in, p->isign, p->stream)) return r;
  LAUNCH(launch_interp(gop(p, p->pt, p->gi), p->grid_in, f, p->stream));
  LAUNCH(launch_cmul_pointwise(p->prec, p->n_trans, p->N, f, p->post, f, p->stream));
  return NUFFTX_OK;
}

}  // namespace

extern "C" {

int nufftx_version(void) { return NUFFTX_VERSION; }

const char* nufftx_last_error(void) { return g_last_error.c_str(); }

int nufftx_kernel_params(double eps, double upsampfac, int* nspread, double* beta, double* c) {
  if (!(eps > 0) || !(upsampfac > 0.5)) return fail(NUFFTX_ERR_INVALID_ARG, "eps must be > 0 and upsampfac > 0.5");
  int ns;
  double b, kc;
  kernel_params(eps, upsampfac, ns, b, kc);
  if (nspread) *nspread = ns;
  if (beta) *beta = b;
  if (c) *c = kc;
  return NUFFTX_OK;
}

int64_t nufftx_grid_size(int64_t n_modes, double upsampfac, int nspread) {
  if (n_modes < 0 || !(upsampfac > 0) || nspread < 1) return -1;
  return grid_size(n_modes, upsampfac, nspread);
}

int nufftx_makeplan(int type, int dim, const int64_t* n_modes, int isign, int64_t n_trans, double eps,
                    double upsampfac, int precision, int modeord, void* stream, nufftx_plan* out) {
  if (!out) return fail(NUFFTX_ERR_INVALID_ARG, "plan output pointer is null");
  *out = nullptr;
  if (int r = check_plan_args(type, dim, n_modes, n_trans, eps, upsampfac, precision, modeord)) return r;

  nufftx_plan p = new nufftx_plan_s();
  p->type = type;
  p->dim = dim;
  p->prec = precision;
  p->isign = isign >= 0 ? 1 : -1;
  p->modeord = modeord;
  p->n_trans = n_trans;
  p->eps = eps;
  p->upsampfac = upsampfac;
  p->stream = static_cast<hipStream_t>(stream);
  p->tune = default_tuning();
  kernel_params(eps, upsampfac, p->ns, p->beta, p->kc);
  p->quad = quad_rule(p->ns, p->beta, p->kc);
  p->g.dim = p->gi.dim = dim;
  p->g.ns = p->gi.ns = p->ns;
  for (int d = 0; d < 3; ++d) {
    p->n_modes[d] = d < dim ? n_modes[d] : 1;
    if (d < dim) {
      const int64_t nf = type == 3 ? n_modes[d] : grid_size(n_modes[d], upsampfac, p->ns);
      if (nf > (1LL << 30)) {
        delete p;
        return fail(NUFFTX_ERR_INVALID_ARG, "fine grid too large");
      }
      p->g.nf[d] = (int)nf;
      p->gi.nf[d] = type == 3 ? (int)grid_size(nf, upsampfac, p->ns) : 1;
    }
  }
  if (type == 3) {
    // any size the reference accepts (nufft3.py:266); supports longer than the grid wrap
    // onto it, and tile cells that alias one grid cell add up in the flush
    for (int d = 0; d < dim; ++d)
      if (p->g.nf[d] < 1) {
        delete p;
        return fail(NUFFTX_ERR_INVALID_ARG, "type-3 grid size must be positive");
      }
  }
  choose_bins(p->g, p->prec, lds_budget_bytes());
  if (type == 3) choose_bins(p->gi, p->prec, lds_budget_bytes());

  int r = NUFFTX_OK;
  char* gbuf = nullptr;
  const size_t gbytes = (size_t)p->g.grid_cells() * p->n_trans * p->csize();
  if ((r = dev_alloc(gbuf, gbytes))) goto fail_out;
  p->grid = gbuf;
  if (type == 1 || type == 2) {
    for (int d = 0; d < dim; ++d) {
      auto ph = phihat(p->g.nf[d], p->quad);
      if ((r = upload_factors(p->fac[d], deconv_factors(ph, p->n_modes[d], modeord), p->prec))) goto fail_out;
    }
    if ((r = fft_make(p->fft, dim, p->g.nf, n_trans, p->prec))) goto fail_out;
    // Fused FFT + deconvolution passes (Stockham, fft2d.hip) in both directions: on the 2D
    // radial 4M workload (1024^2 grid) nufft2d1 0.198 -> 0.187 ms against rocFFT + the shuffle
    // kernel, nufft2d2 0.160 -> 0.156 ms against the earlier four-step passes.  NUFFTX_ROCFFT
    // takes rocFFT + the separate deconvolution kernels in both directions.
    if (fft2_applicable(dargs(p, false)) && !std::getenv("NUFFTX_ROCFFT")) {
      if (fft2_plan_init(p->fft2, dargs(p, false)) != hipSuccess) {
        r = fail(NUFFTX_ERR_ALLOC, "fused FFT twiddle tables and intermediate buffer");
        goto fail_out;
      }
      p->fft2_pad = true;
      p->fft2_shuffle = true;
    }
  } else {
    char* gi = nullptr;
    const size_t gib = (size_t)p->gi.grid_cells() * p->n_trans * p->csize();
    if ((r = dev_alloc(gi, gib))) goto fail_out;
    p->grid_in = gi;
    for (int d = 0; d < dim; ++d) {
      auto ph = phihat(p->gi.nf[d], p->quad);
      if ((r = upload_factors(p->fac_in[d], deconv_factors(ph, p->g.nf[d], 0), p->prec))) goto fail_out;
    }
    if ((r = fft_make(p->fft_in, dim, p->gi.nf, n_trans, p->prec))) goto fail_out;
    if ((r = dev_alloc(p->t3, 1))) goto fail_out;
    if ((r = dev_alloc(p->red, 12 * 256))) goto fail_out;
  }
  *out = p;
  return NUFFTX_OK;
fail_out:
  plan_free(p);
  delete p;
  return r;
}

int nufftx_makeplan_grid(int dim, const int64_t* nf, int nspread, double beta, double kc, int64_t n_trans,
                         int precision, void* stream, nufftx_plan* out) {
  if (!out) return fail(NUFFTX_ERR_INVALID_ARG, "plan output pointer is null");
  *out = nullptr;
  if (int r = check_grid_args(dim, nf, nspread, n_trans, precision)) return r;
  nufftx_plan p = new nufftx_plan_s();
  p->type = 0;
  p->dim = dim;
  p->prec = precision;
  p->n_trans = n_trans;
  p->ns = nspread;
  p->beta = beta;
  p->kc = kc;
  p->stream = static_cast<hipStream_t>(stream);
  p->tune = default_tuning();
  p->g.dim = dim;
  p->g.ns = nspread;
  for (int d = 0; d < 3; ++d) p->g.nf[d] = d < dim ? (int)nf[d] : 1;
  choose_bins(p->g, p->prec, lds_budget_bytes());
  *out = p;
  return NUFFTX_OK;
}

int64_t nufftx_plan_grid_size(nufftx_plan p, int d, int inner) {
  if (!p || d < 0 || d > 2) return -1;
  return inner ? p->gi.nf[d] : p->g.nf[d];
}

int nufftx_set_stream(nufftx_plan p, void* stream) {
  if (int r = check_plan(p)) return r;
  if (p->cached) return fail(NUFFTX_ERR_STATE, "a cached plan keeps the stream it was acquired for");
  p->stream = static_cast<hipStream_t>(stream);
  return NUFFTX_OK;
}

int nufftx_setpts(nufftx_plan p, int64_t M, const void* x, const void* y, const void* z, int64_t N, const void* s,
                  const void* t, const void* u) {
  if (int r = check_plan(p)) return r;
  if (M < 0 || M > (int64_t)INT32_MAX - 1) return fail(NUFFTX_ERR_INVALID_ARG, "M out of range");
  const void* xs[3] = {x, y, z};
  for (int d = 0; d < p->dim; ++d)
    if (M > 0 && !xs[d]) return fail(NUFFTX_ERR_INVALID_ARG, "missing coordinate array");
  p->pts_set = false;
  p->M = M;
  if (p->type != 3) {
    p->ps.M = M;
    for (int d = 0; d < 3; ++d) p->ps.coord[d] = d < p->dim ? xs[d] : nullptr;
    p->ps.coord_f64 = p->prec == kF64;
    if (int r = sort_points(p, p->ps, p->g, p->type != 1, p->type != 2)) return r;
    p->pts_set = true;
    return NUFFTX_OK;
  }
  if (N < 0 || N > (int64_t)INT32_MAX - 1) return fail(NUFFTX_ERR_INVALID_ARG, "N out of range");
  const void* ss[3] = {s, t, u};
  for (int d = 0; d < p->dim; ++d)
    if (N > 0 && !ss[d]) return fail(NUFFTX_ERR_INVALID_ARG, "missing target array");
  p->N = N;
  int r;
  for (int d = 0; d < p->dim; ++d) {
    if ((r = dev_grow(p->xr[d], p->cap_xr[d], M, p->retired))) return r;
    if ((r = dev_grow(p->sr[d], p->cap_sr[d], N, p->retired))) return r;
  }
  if ((r = dev_grow(p->pre, p->cap_pre, 2 * M, p->retired))) return r;
  if ((r = dev_grow(p->post, p->cap_post, 2 * N, p->retired))) return r;
  {
    char* cb = static_cast<char*>(p->cbuf);
    if ((r = dev_grow(cb, p->cap_cbuf, M * p->n_trans * (int64_t)p->csize(), p->retired))) return r;
    p->cbuf = cb;
  }
  const void* const xa[3] = {x, p->dim > 1 ? y : x, p->dim > 2 ? z : x};
  const void* const sa[3] = {s, p->dim > 1 ? t : s, p->dim > 2 ? u : s};
  LAUNCH(launch_t3_setup(p->dim, p->prec, M, xa, N, sa, p->g.nf, p->upsampfac, p->t3, p->red, 12 * 256, p->stream));
  double* xr3[3] = {p->xr[0], p->xr[1] ? p->xr[1] : p->xr[0], p->xr[2] ? p->xr[2] : p->xr[0]};
  double* sr3[3] = {p->sr[0], p->sr[1] ? p->sr[1] : p->sr[0], p->sr[2] ? p->sr[2] : p->sr[0]};
  LAUNCH(launch_t3_prep_src(p->dim, p->prec, M, xa, p->t3, p->isign, xr3, p->pre, p->stream));
  LAUNCH(launch_t3_prep_tgt(p->dim, p->prec, N, sa, p->t3, p->isign, p->quad, sr3, p->post, p->stream));
  p->ps.M = M;
  p->pt.M = N;
  for (int d = 0; d < 3; ++d) {
    p->ps.coord[d] = d < p->dim ? p->xr[d] : nullptr;
    p->pt.coord[d] = d < p->dim ? p->sr[d] : nullptr;
  }
  p->ps.coord_f64 = p->pt.coord_f64 = true;
  if ((r = sort_points(p, p->ps, p->g, false, true))) return r;
  if ((r = sort_points(p, p->pt, p->gi, true, false))) return r;
  p->pts_set = true;
  return NUFFTX_OK;
}

int nufftx_execute_sign(nufftx_plan p, void* c, void* f, int isign) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  if (p->type == 0) return fail(NUFFTX_ERR_STATE, "grid plans have no execute; use spread/interp");
  if ((!c && p->M > 0) || !f) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  const int s = isign >= 0 ? 1 : -1;
  if (p->type == 1) return exec_type1(p, c, f, s);
  if (p->type == 2) return exec_type2(p, c, f, s);
  if (s != p->isign) return fail(NUFFTX_ERR_INVALID_ARG, "type-3 plans run with their planned isign");
  return exec_type3(p, c, f);
}

int nufftx_execute(nufftx_plan p, void* c, void* f) {
  if (int r = check_plan(p)) return r;
  return nufftx_execute_sign(p, c, f, p->isign);
}

int nufftx_execute_many(int count, const nufftx_plan* plans, void* const* c, void* const* f) {
  if (count < 1 || !plans || !c || !f) return fail(NUFFTX_ERR_INVALID_ARG, "execute_many: empty or null batch");
  for (int i = 0; i < count; ++i) {
    if (int r = check_plan(plans[i])) return r;
    if (plans[i]->type != plans[0]->type || (plans[i]->type != 1 && plans[i]->type != 2))
      return fail(NUFFTX_ERR_INVALID_ARG, "execute_many: all plans must be type 1, or all type 2");
    if (!plans[i]->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
    if ((!c[i] && plans[i]->M > 0) || !f[i]) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
    for (int k = 0; k < i; ++k)
      if (plans[k] == plans[i]) return fail(NUFFTX_ERR_INVALID_ARG, "execute_many: a plan appears twice");
  }
  bool shared = count > 1 && count <= kMaxManyPlans;
  for (int i = 1; i < count && shared; ++i) shared = plans[i]->stream == plans[0]->stream;
  if (!shared) {
    for (int i = 0; i < count; ++i)
      if (int r = nufftx_execute(plans[i], c[i], f[i])) return r;
    return NUFFTX_OK;
  }
  const hipStream_t st = plans[0]->stream;
  GridOpArgs gas[kMaxManyPlans];
  void* grids[kMaxManyPlans];
  for (int i = 0; i < count; ++i) {
    gas[i] = gop(plans[i], plans[i]->ps, plans[i]->g);
    grids[i] = plans[i]->grid;
  }
  if (plans[0]->type == 1) {
    for (int i = 0; i < count; ++i) {
      const nufftx_plan p = plans[i];
      LAUNCH(launch_zero(p->grid, p->g.grid_cells() * p->n_trans * (int64_t)p->csize(), st));
    }
    hipError_t e = launch_spread_many(gas, count, c, grids, st);
    if (e == hipErrorNotSupported) {
      (void)hipGetLastError();
      for (int i = 0; i < count; ++i) LAUNCH(launch_spread(gas[i], c[i], grids[i], st));
    } else {
      LAUNCH(e);
    }
    for (int i = 0; i < count; ++i)
      if (int r = grid_to_modes(plans[i], f[i], plans[i]->isign)) return r;
    return NUFFTX_OK;
  }
  for (int i = 0; i < count; ++i)
    if (int r = modes_to_grid(plans[i], f[i], plans[i]->isign)) return r;
  hipError_t e = launch_interp_many(gas, count, grids, c, st);
  if (e == hipErrorNotSupported) {
    (void)hipGetLastError();
    for (int i = 0; i < count; ++i) LAUNCH(launch_interp(gas[i], grids[i], c[i], st));
  } else {
    LAUNCH(e);
  }
  return NUFFTX_OK;
}

int nufftx_vjp(nufftx_plan p, const void* src, const void* g, void* grad_src, void* gx, void* gy, void* gz) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  if (p->type != 1 && p->type != 2) return fail(NUFFTX_ERR_INVALID_ARG, "vjp is defined for type-1/2 plans");
  if (!g) return fail(NUFFTX_ERR_INVALID_ARG, "null cotangent");
  void* gc[3] = {gx, p->dim > 1 ? gy : nullptr, p->dim > 2 ? gz : nullptr};
  const bool want_coord = gc[0] || gc[1] || gc[2];
  const size_t rsz = p->prec == kF32 ? 4 : 8;
  if (want_coord && !src) return fail(NUFFTX_ERR_INVALID_ARG, "coordinate gradients need the primal input");
  for (int d = 0; d < 3; ++d)
    if (gc[d]) LAUNCH(launch_zero(gc[d], p->M * (int64_t)rsz, p->stream));
  const int s = -p->isign;
  if (p->type == 1) {
    // G = FFT_{-isign}(pad(g / phihat)) -- adjoint of deconvolve-shuffle and FFT
    if (int r = modes_to_grid(p, g, s)) return r;
    if (want_coord) {
      LAUNCH(launch_interp_grad(gop(p, p->ps, p->g), p->grid, src, grad_src, gc, p->stream));
    } else if (grad_src) {
      LAUNCH(launch_interp(gop(p, p->ps, p->g), p->grid, grad_src, p->stream));
    }
    return NUFFTX_OK;
  }
  // type 2
  if (want_coord) {
    if (int r = modes_to_grid(p, src, p->isign)) return r;
    LAUNCH(launch_interp_grad(gop(p, p->ps, p->g), p->grid, g, nullptr, gc, p->stream));
  }
  if (grad_src) return exec_type1(p, g, grad_src, s);
  return NUFFTX_OK;
}

int nufftx_destroy(nufftx_plan p) {
  if (!p) return NUFFTX_OK;
  if (p->cached) return fail(NUFFTX_ERR_STATE, "a cached plan is returned with nufftx_plan_release");
  plan_free(p);
  delete p;
  return NUFFTX_OK;
}

int nufftx_spread(nufftx_plan p, const void* c, void* fw) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  if (!fw || (!c && p->M > 0)) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  GridOpArgs ga = gop(p, p->ps, p->g);
  ga.zero_out = true;
  LAUNCH(launch_spread(ga, c, fw, p->stream));
  return NUFFTX_OK;
}

int nufftx_spread_add(nufftx_plan p, const void* c, void* fw) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  if (!fw || (!c && p->M > 0)) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  LAUNCH(launch_spread(gop(p, p->ps, p->g), c, fw, p->stream));
  return NUFFTX_OK;
}

int nufftx_interp(nufftx_plan p, const void* fw, void* c) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  if (!fw || (!c && p->M > 0)) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  LAUNCH(launch_interp(gop(p, p->ps, p->g), fw, c, p->stream));
  return NUFFTX_OK;
}

int nufftx_gather_grad(nufftx_plan p, const void* G, const void* P, void* gval, void* gx, void* gy, void* gz) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  void* gc[3] = {gx, p->dim > 1 ? gy : nullptr, p->dim > 2 ? gz : nullptr};
  const bool want_coord = gc[0] || gc[1] || gc[2];
  if (!G || (want_coord && !P)) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  const size_t rsz = p->prec == kF32 ? 4 : 8;
  for (int d = 0; d < 3; ++d)
    if (gc[d]) LAUNCH(launch_zero(gc[d], p->M * (int64_t)rsz, p->stream));
  if (want_coord) {
    LAUNCH(launch_interp_grad(gop(p, p->ps, p->g), G, P, gval, gc, p->stream));
  } else if (gval) {
    LAUNCH(launch_interp(gop(p, p->ps, p->g), G, gval, p->stream));
  }
  return NUFFTX_OK;
}

int nufftx_get_sort_info(nufftx_plan p, int inner, int* nbins, int* bin_size, int* nbd) {
  if (int r = check_plan(p)) return r;
  const Geometry& g = inner ? p->gi : p->g;
  if (nbins) *nbins = g.nbins;
  for (int d = 0; d < 3; ++d) {
    if (bin_size) bin_size[d] = g.bs[d];
    if (nbd) nbd[d] = g.nb[d];
  }
  return NUFFTX_OK;
}

int nufftx_copy_sort(nufftx_plan p, int inner, int32_t* perm, int32_t* bin_start) {
  if (int r = check_plan(p)) return r;
  if (!p->pts_set) return fail(NUFFTX_ERR_STATE, "setpts has not been called");
  const PointSet& ps = inner ? p->pt : p->ps;
  const Geometry& g = inner ? p->gi : p->g;
  if (perm && ps.M > 0) HIPCHK(hipMemcpyAsync(perm, ps.perm, ps.M * sizeof(int32_t), hipMemcpyDeviceToDevice, p->stream));
  if (bin_start)
    HIPCHK(hipMemcpyAsync(bin_start, ps.bin_start, (g.nbins + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, p->stream));
  return NUFFTX_OK;
}

int nufftx_kernel_fourier_series(int64_t nf, int nspread, double beta, double c, double* out) {
  if (nf < 1 || nspread < 2 || nspread > 16 || !out) return fail(NUFFTX_ERR_INVALID_ARG, "bad kernel Fourier series arguments");
  const std::vector<double> ph = phihat(nf, quad_rule(nspread, beta, c));
  std::copy(ph.begin(), ph.end(), out);
  return NUFFTX_OK;
}

int nufftx_deconvolve(int direction, int dim, const int64_t* n_modes, const int64_t* nf, int64_t n_trans,
                      const void* phihat1, const void* phihat2, const void* phihat3, int modeord, int precision,
                      void* stream, const void* in, void* out) {
  if (direction != 0 && direction != 1) return fail(NUFFTX_ERR_INVALID_ARG, "direction must be 0 (shuffle) or 1 (pad)");
  if (dim < 1 || dim > 3 || !n_modes || !nf) return fail(NUFFTX_ERR_INVALID_ARG, "dim must be 1, 2 or 3");
  if (precision != NUFFTX_F32 && precision != NUFFTX_F64) return fail(NUFFTX_ERR_INVALID_ARG, "bad precision");
  if (modeord != 0 && modeord != 1) return fail(NUFFTX_ERR_INVALID_ARG, "modeord must be 0 or 1");
  if (n_trans < 0) return fail(NUFFTX_ERR_INVALID_ARG, "n_trans must be >= 0");
  const void* ph[3] = {phihat1, phihat2, phihat3};
  DeconvArgs a;
  a.dim = dim;
  a.prec = precision == NUFFTX_F32 ? kF32 : kF64;
  a.n_trans = n_trans;
  a.modeord = modeord;
  a.from_phihat = true;
  for (int d = 0; d < 3; ++d) {
    a.nf[d] = 1;
    a.nm[d] = 1;
    a.fac[d] = nullptr;
    if (d >= dim) continue;
    if (n_modes[d] < 1 || nf[d] < n_modes[d] || nf[d] > INT32_MAX)
      return fail(NUFFTX_ERR_INVALID_ARG, "need 1 <= n_modes <= nf < 2^31 on every axis");
    if (!ph[d]) return fail(NUFFTX_ERR_INVALID_ARG, "null phihat pointer");
    a.nf[d] = (int)nf[d];
    a.nm[d] = (int)n_modes[d];
    a.fac[d] = ph[d];
  }
  if (n_trans == 0) return NUFFTX_OK;
  if (!in || !out) return fail(NUFFTX_ERR_INVALID_ARG, "null data pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (direction == 0)
    LAUNCH(launch_deconv_shuffle(a, in, out, s));
  else
    LAUNCH(launch_deconv_pad(a, in, out, s));
  return NUFFTX_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ plan cache
// Plans keyed by their parameters, the stream and the device.  A plan is held by one caller
// between acquire and release; release returns it to the idle pool, where the next acquire
// with the same key takes it.  Work that a released plan still has in flight is ordered
// before the next holder's by the key's stream, so release does not wait for the GPU.  At
// most kCacheIdleMax idle plans are kept: beyond that the least recently released one is
// destroyed (after a synchronisation of its device, as its buffers are freed).
namespace {

struct CacheKey {
  int kind = 0;  // 0 transform (nufftx_makeplan), 1 grid operator (nufftx_makeplan_grid)
  int type = 0, dim = 0, isign = 0, prec = 0, modeord = 0, ns = 0;
  int64_t n[3] = {0, 0, 0};  // n_modes (transform) or nf (grid)
  int64_t n_trans = 0;
  double eps = 0, upsampfac = 0, beta = 0, kc = 0;
  void* stream = nullptr;
  int device = -1;
  bool operator==(const CacheKey& o) const {
    return kind == o.kind && type == o.type && dim == o.dim && isign == o.isign && prec == o.prec &&
           modeord == o.modeord && ns == o.ns && n[0] == o.n[0] && n[1] == o.n[1] && n[2] == o.n[2] &&
           n_trans == o.n_trans && eps == o.eps && upsampfac == o.upsampfac && beta == o.beta && kc == o.kc &&
           stream == o.stream && device == o.device;
  }
};

struct CacheEntry {
  CacheKey key;
  nufftx_plan plan = nullptr;
  bool busy = false;
  uint64_t released = 0;  // release tick, for least-recently-released eviction
};

constexpr int kCacheIdleMax = 32;

struct PlanCache {
  std::mutex mu;
  std::vector<CacheEntry> entries;
  uint64_t tick = 0;
  int64_t made = 0, reused = 0;
};

PlanCache& plan_cache() {
  static PlanCache* c = new PlanCache();  // never destroyed: plans may be released at exit
  return *c;
}

int current_device() {
  int d = -1;
  if (hipGetDevice(&d) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  return d;
}

// an idle plan of this key, marked busy; nullptr when there is none
nufftx_plan cache_take(const CacheKey& k) {
  PlanCache& pc = plan_cache();
  std::lock_guard<std::mutex> lk(pc.mu);
  for (CacheEntry& e : pc.entries)
    if (!e.busy && e.key == k) {
      e.busy = true;
      ++pc.reused;
      return e.plan;
    }
  return nullptr;
}

void cache_insert_busy(const CacheKey& k, nufftx_plan p) {
  PlanCache& pc = plan_cache();
  p->cached = true;
  std::lock_guard<std::mutex> lk(pc.mu);
  CacheEntry e;
  e.key = k;
  e.plan = p;
  e.busy = true;
  pc.entries.push_back(e);
  ++pc.made;
}

// Destroy cached plans taken out of the pool (outside the lock).  Their streams may have been
// destroyed since, so the wait for their work in flight is a synchronisation of their device.
void cache_destroy(std::vector<CacheEntry>& out) {
  int cur = -1;
  const bool have_cur = hipGetDevice(&cur) == hipSuccess;
  for (CacheEntry& e : out) {
    if (e.key.device >= 0) (void)hipSetDevice(e.key.device);
    (void)hipDeviceSynchronize();
    e.plan->cached = false;
    (void)nufftx_destroy(e.plan);
  }
  if (have_cur && !out.empty()) (void)hipSetDevice(cur);
  (void)hipGetLastError();
  out.clear();
}

}  // namespace

extern "C" {

int nufftx_plan_acquire(int type, int dim, const int64_t* n_modes, int isign, int64_t n_trans, double eps,
                        double upsampfac, int precision, int modeord, void* stream, nufftx_plan* out) {
  if (!out) return fail(NUFFTX_ERR_INVALID_ARG, "plan output pointer is null");
  *out = nullptr;
  if (int r = check_plan_args(type, dim, n_modes, n_trans, eps, upsampfac, precision, modeord)) return r;
  CacheKey k;
  k.kind = 0;
  k.type = type;
  k.dim = dim;
  // type 1/2 plans serve either sign (the sign is set on the plan at acquire); type 3 bakes it in
  k.isign = type == 3 ? (isign >= 0 ? 1 : -1) : 0;
  k.prec = precision;
  k.modeord = modeord;
  for (int d = 0; d < dim; ++d) k.n[d] = n_modes[d];
  k.n_trans = n_trans;
  k.eps = eps;
  k.upsampfac = upsampfac;
  k.stream = stream;
  k.device = current_device();
  nufftx_plan p = cache_take(k);
  if (!p) {
    if (int r = nufftx_makeplan(type, dim, n_modes, isign, n_trans, eps, upsampfac, precision, modeord, stream, &p))
      return r;
    cache_insert_busy(k, p);
  }
  p->isign = isign >= 0 ? 1 : -1;
  *out = p;
  return NUFFTX_OK;
}

int nufftx_plan_acquire_grid(int dim, const int64_t* nf, int nspread, double beta, double kc, int64_t n_trans,
                             int precision, void* stream, nufftx_plan* out) {
  if (!out) return fail(NUFFTX_ERR_INVALID_ARG, "plan output pointer is null");
  *out = nullptr;
  if (int r = check_grid_args(dim, nf, nspread, n_trans, precision)) return r;
  CacheKey k;
  k.kind = 1;
  k.dim = dim;
  k.prec = precision;
  k.ns = nspread;
  for (int d = 0; d < dim; ++d) k.n[d] = nf[d];
  k.n_trans = n_trans;
  k.beta = beta;
  k.kc = kc;
  k.stream = stream;
  k.device = current_device();
  nufftx_plan p = cache_take(k);
  if (!p) {
    if (int r = nufftx_makeplan_grid(dim, nf, nspread, beta, kc, n_trans, precision, stream, &p)) return r;
    cache_insert_busy(k, p);
  }
  *out = p;
  return NUFFTX_OK;
}

int nufftx_plan_release(nufftx_plan p) {
  if (!p) return fail(NUFFTX_ERR_INVALID_ARG, "null plan");
  PlanCache& pc = plan_cache();
  std::vector<CacheEntry> evict;
  {
    std::lock_guard<std::mutex> lk(pc.mu);
    CacheEntry* mine = nullptr;
    for (CacheEntry& e : pc.entries)
      if (e.plan == p) mine = &e;
    if (!mine) return fail(NUFFTX_ERR_INVALID_ARG, "plan was not acquired from the plan cache");
    if (!mine->busy) return fail(NUFFTX_ERR_STATE, "plan released twice");
    mine->busy = false;
    mine->released = ++pc.tick;
    int idle = 0;
    for (const CacheEntry& e : pc.entries) idle += e.busy ? 0 : 1;
    while (idle > kCacheIdleMax) {
      size_t lru = pc.entries.size();
      for (size_t i = 0; i < pc.entries.size(); ++i)
        if (!pc.entries[i].busy && (lru == pc.entries.size() || pc.entries[i].released < pc.entries[lru].released))
          lru = i;
      evict.push_back(pc.entries[lru]);
      pc.entries.erase(pc.entries.begin() + (std::ptrdiff_t)lru);
      --idle;
    }
  }
  cache_destroy(evict);
  return NUFFTX_OK;
}

int nufftx_plan_cache_clear(void) {
  PlanCache& pc = plan_cache();
  std::vector<CacheEntry> idle;
  {
    std::lock_guard<std::mutex> lk(pc.mu);
    std::vector<CacheEntry> keep;
    for (CacheEntry& e : pc.entries) {
      if (e.busy)
        keep.push_back(e);
      else
        idle.push_back(e);
    }
    pc.entries.swap(keep);
  }
  const int n = (int)idle.size();
  cache_destroy(idle);
  return n;
}

int nufftx_plan_cache_stats(int64_t* idle, int64_t* held, int64_t* made, int64_t* reused) {
  PlanCache& pc = plan_cache();
  std::lock_guard<std::mutex> lk(pc.mu);
  int64_t ni = 0, nb = 0;
  for (const CacheEntry& e : pc.entries) (e.busy ? nb : ni) += 1;
  if (idle) *idle = ni;
  if (held) *held = nb;
  if (made) *made = pc.made;
  if (reused) *reused = pc.reused;
  return NUFFTX_OK;
}

int64_t nufftx_device_allocations(void) { return g_dev_allocs.load(std::memory_order_relaxed); }

}  // extern "C"
