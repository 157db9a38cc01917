// Internal interfaces between the nufftx translation units (host side).
#pragma once
#include <vector>

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <string>

#include "common.hpp"

namespace nufftx {

enum Prec { kF32 = 0, kF64 = 1 };

// Records msg as the calling thread's last error (nufftx_last_error) and returns code.
int fail(int code, const std::string& msg);

// Binning of one fine grid: bins of bs[d] cells, tiles of bs[d] + ns cells.
struct Geometry {
  int dim = 1;
  int nf[3] = {1, 1, 1};
  int bs[3] = {1, 1, 1};
  int nb[3] = {1, 1, 1};
  int nbins = 1;
  int ns = 2;
  int tile[3] = {1, 1, 1};  // bs + ns per used dim, 1 otherwise
  int64_t tile_cells = 1;
  int64_t grid_cells() const { return (int64_t)nf[0] * nf[1] * nf[2]; }
};

// per-block partial maxima of the strength bound (k_absmax_part): at most this many per transform
constexpr int kAbsmaxBlocks = 1024;

// A sorted set of points on one geometry.  Coordinates are not copied: the
// kernels re-read them through perm and refold them in FP64.
struct PointSet {
  int64_t M = 0;
  const void* coord[3] = {nullptr, nullptr, nullptr};
  bool coord_f64 = false;
  // sort state
  int NB = 0;        // sort chunks (rows of the histogram matrix)
  int CH = 0;        // points per chunk
  int maxsub = 0;    // max points per subproblem
  int max_subprob = 0;
  // device buffers (owned)
  int32_t* perm = nullptr;       // M
  int32_t* keys = nullptr;       // M, bin key of every point (index order)
  int32_t* hist = nullptr;       // nbins * NB (+ scan scratch)
  int32_t* scan_part = nullptr;  // scan partials
  int32_t* scan_sum = nullptr;   // segment sums of the histogram (zero between sorts)
  int64_t cap_sum = 0;
  int32_t* bin_start = nullptr;  // nbins + 1
  int32_t* sub_start = nullptr;  // nbins + 1, first subproblem of every bin
  int4* subprob = nullptr;       // max_subprob
  int32_t* nsub = nullptr;       // 1
  // 2D FP32 point sets: point records in sorted order (sort.hip), optionally re-ordered
  // by sub-bin inside every subproblem for the matrix-core kernels (prep.hip)
  bool has_recs = false;
  bool recs_subbin = false;
  float4* recs = nullptr;        // M full records {t_x, t_y, meta, p} (matrix-core kernels)
  bool has_crec = false;
  uint2* crec = nullptr;         // M compact records (LDS-tiled kernels)
  int64_t cap_crec = 0;
  int32_t* sub_off = nullptr;    // max_subprob * 17 sub-bin starts per subproblem
  float* cmax = nullptr;          // n_trans: max |re|, |im| of the strengths, then the per-block partials
  // 2D FP32 point sets spread by the matrix-core window kernel: the inverse permutation
  // (slot[i] = sorted position of point i, written by the sort) and, per execute, every
  // transform's points packed in sorted order as {x word, y word, re c, im c} (compact
  // words as in crec), so that the spread reads its points contiguously
  int32_t* slot = nullptr;       // M
  bool has_slot = false;
  uint4* packed = nullptr;       // n_trans * M (2D); M {x, y, z word, 0} (3D)
  float2* cs = nullptr;          // 3D: the strengths of every transform in sorted order, n_trans * M
  int64_t cap_slot = 0, cap_packed = 0, cap_cs = 0;
  int64_t cap_recs = 0, cap_sub_off = 0, cap_cmax = 0;
  int64_t cap_M = 0, cap_keys = 0, cap_hist = 0, cap_part = 0, cap_bins = 0, cap_sub = 0, cap_sub_start = 0;
  std::vector<void*> retired;    // outgrown buffers, freed by pointset_free
};

struct SortTuning {
  int chunk = 8192;            // points per histogram row (sort chunk)
  int64_t hist_cap = 8 << 20;  // max entries of the histogram matrix
  int maxsub = 1024;           // points per subproblem
};

// Whether a geometry takes the matrix-core window spread (si_win.hpp): 2D FP32, ns <= 8
// (window 16 = SB + ns with SB >= 8), square bins of 4 x 4 sub-bins, subproblems of at most
// 4096 points (eight per thread of a 512-thread block)
constexpr int kWinMaxSub = 4096;
// Whether a geometry takes the 3D matrix-core window spread (si_win3.hpp): 3D FP32, ns = 8
// (its windows cost the same at every ns, and the bank-spread kernel's ns^3 atomics per point
// are cheaper below 8: 1.06 against 1.48 ms for 8M clustered points at ns = 6), bins of whole
// 8 x 8 x 4 sub-bins (at most 16 of them), tiles of at most 127 cells a side, subproblems of
// at most 2048 points (four per thread of a 512-thread block)
constexpr int kWin3MaxSub = 2048;
inline bool win3d_geometry(const Geometry& g, int prec, int maxsub) {
  if (g.dim != 3 || prec != 0 || g.ns != 8 || maxsub > kWin3MaxSub) return false;
  if (g.bs[0] % 8 || g.bs[1] % 8 || g.bs[2] % 4) return false;
  if ((g.bs[0] / 8) * (g.bs[1] / 8) * (g.bs[2] / 4) > 16) return false;
  for (int d = 0; d < 3; ++d)
    if (g.tile[d] > 127 || g.tile[d] != g.bs[d] + g.ns) return false;
  return true;
}
// The window spread is the default where it applies; NUFFTX_SPREAD_WIN=0 selects the
// LDS-atomic bank-spread kernel instead (read when points are set and per call; measured on
// the 2D radial workload, 4M points, 1024^2 grid: 143 us for packing + the fp16-split window
// kernel against 200-213 us)
inline bool win_spread_enabled() {
  const char* e = std::getenv("NUFFTX_SPREAD_WIN");
  return e == nullptr || e[0] != '0';
}
// NUFFTX_INTERP_WIN=1 turns the window interpolation on (read when points are set and per
// call).  Off by default: on the 2D radial workload it took 113-126 us against 73 us for the
// wide gather (its 16-point groups run a long dependent chain per wave: weights, LDS rows,
// split, three MFMAs, dot product, two cross-lane adds).
inline bool win_interp_enabled() {
  const char* e = std::getenv("NUFFTX_INTERP_WIN");
  return e != nullptr && e[0] == '1';
}
inline bool win2d_geometry(const Geometry& g, int prec, int maxsub) {
  if (g.dim != 2 || prec != 0 || g.ns > 8 || maxsub > kWinMaxSub) return false;
  const int bs = 4 * (16 - g.ns);
  return g.bs[0] == bs && g.bs[1] == bs && g.tile[0] == bs + g.ns && g.tile[1] == bs + g.ns;
}

// Tile size / LDS budget helpers
void choose_bins(Geometry& g, int prec, int lds_budget);
int lds_budget_bytes();
int lds_device_max_bytes();

// ---- sort.hip
// rec_kind (2D only): also write point records, 1 = full (matrix-core kernels),
// 2 = compact (LDS-tiled kernels); sub_sb: sub-bin side of the matrix-core
// geometry (16 - ns) to record sub-bins in full records, 0 otherwise
// Sizes of a point set's sort state (pointset_sort): waves per chunk, points per wave, 16-bit
// counters, chunks (histogram rows), histogram entries, segment partials, subproblem capacity.
struct SortSizes {
  int W = 1;
  int64_t SC = 64;
  bool c16 = false;
  int64_t NB = 1, nhist = 0, nparts = 0, max_sub = 0;
};
SortSizes sort_sizes(int64_t M, const Geometry& g, const SortTuning& tune, int rec_kind);

// want_slot (2D, 3D): also write the inverse permutation, and size the packed records for
// n_trans transforms (n_trans 0: left to the first spread)
hipError_t pointset_sort(PointSet& ps, const Geometry& g, const SortTuning& tune, hipStream_t s, std::string& err,
                         int rec_kind = 0, int sub_sb = 0, bool want_slot = false, int64_t n_trans = 1);
// packed[t][slot[i]] = {x word, y word, c[t][i]} of a 2D point set sorted with want_slot (FP32
// strengths; a non-finite coordinate makes its strength NaN); the same launch clears
// zero_bytes (a multiple of 16) at zero, when given
hipError_t pointset_pack2d(PointSet& ps, const Geometry& g, const void* c, int64_t n_trans, hipStream_t s,
                           void* zero = nullptr, int64_t zero_bytes = 0);
// 3D: packed[slot[i]] = {x word, y word, z word, 0}, cs[t][slot[i]] = c[t][i] (NaN for a
// non-finite coordinate); the same launch clears zero_bytes (a multiple of 16) at zero
hipError_t pointset_pack3d(PointSet& ps, const Geometry& g, const void* c, int64_t n_trans, hipStream_t s,
                           void* zero = nullptr, int64_t zero_bytes = 0);

// Device buffer of at least `need` elements (contents not kept when it grows).  The
// outgrown buffer is not freed here: hipFree waits for the whole device, which would
// stall every other stream (a serving loop runs one plan per stream).  It is moved to
// `retired` and freed when its owner is destroyed; capacity grows by at least half
// each time, so the retired buffers stay below twice the live one.  On failure the old
// buffer and capacity are kept.
// Every device allocation of the library goes through dev_malloc, which counts them
// (nufftx_device_allocations reports the count).
hipError_t dev_malloc(void** p, size_t bytes);

template <typename T>
hipError_t grow_buffer(T*& p, int64_t& cap, int64_t need, std::vector<void*>& retired) {
  if (need <= cap && p) return hipSuccess;
  int64_t n = need > 1 ? need : 1;
  if (p && cap + cap / 2 > n) n = cap + cap / 2;
  T* q = nullptr;
  hipError_t e = dev_malloc((void**)&q, (size_t)n * sizeof(T));
  if (e != hipSuccess) return e;
  if (p) retired.push_back(p);
  p = q;
  cap = n;
  return hipSuccess;
}

// ---- prep.hip: sub-bin order and point records for the 2D FP32 matrix-core kernels
bool subsort_applicable(const SortTuning& tune, const Geometry& g, int prec);
hipError_t pointset_subsort(PointSet& ps, const Geometry& g, hipStream_t s);
void pointset_free(PointSet& ps);

// ---- spreadinterp.hip
// The polynomial ES kernel for (ns, beta, c) (common.hpp), fitted once per parameter set and
// kept for the process; on = 0 where ns is not kEsPolyNs or the fit misses kEsPolyTol.
// NUFFTX_ES_POLY=0 (read per call) keeps the exp form everywhere.
EsPoly es_poly_table(int ns, double beta, double c);

struct GridOpArgs {
  const PointSet* ps;
  const Geometry* g;
  int prec;
  int64_t n_trans;
  double beta, c;
  double tol = 0.0;  // the plan's requested accuracy (0: unknown), selects packed FP32 tiles
  bool zero_out = false;  // launch_spread overwrites fw instead of adding to it
};
// fw[n_trans][grid] += spread(c[n_trans][M]) (fw = ... with zero_out: the clear then rides on
// the window path's packing pass, or is a fill kernel of its own on the other paths).
hipError_t launch_spread(const GridOpArgs& a, const void* c, void* fw, hipStream_t s);
// c[n_trans][M] = interp(fw)
hipError_t launch_interp(const GridOpArgs& a, const void* fw, void* c, hipStream_t s);
constexpr int kMaxManyPlans = 8;
// The same for up to kMaxManyPlans point sets in one launch (FP32, 2D/3D, one kernel size and n_trans;
// points and grids may differ).  hipErrorNotSupported: launch them one by one instead.
hipError_t launch_spread_many(const GridOpArgs* a, int count, const void* const* c, void* const* fw, hipStream_t s);
hipError_t launch_interp_many(const GridOpArgs* a, int count, const void* const* fw, void* const* c, hipStream_t s);
// Fused gather with kernel-derivative weights.  For each point i and transform t:
//   S0 = sum G w, S_ax = sum G dW/d ax   (dW/d ax includes the -nf/2pi chain factor)
//   gval[t][i] = S0                               (if gval != nullptr)
//   gcoord_ax[i] += Re(conj(P[t][i]) * S_ax)      (if gcoord[ax] != nullptr; zeroed by caller)
hipError_t launch_interp_grad(const GridOpArgs& a, const void* G, const void* P, void* gval, void* const gcoord[3],
                              hipStream_t s);

// ---- elementwise.hip
hipError_t launch_zero(void* p, int64_t bytes, hipStream_t s);
// Type-1 mode extraction (deconvolve.py:81-366) and type-2 padding (deconvolve.py:114-468).
struct DeconvArgs {
  int dim;
  int prec;
  int64_t n_trans;
  int nf[3];
  int nm[3];
  int modeord;
  const void* fac[3];  // per dim, T, in output mode order; from_phihat: phihat[0 .. n/2] instead
  bool from_phihat = false;
};
hipError_t launch_deconv_shuffle(const DeconvArgs& a, const void* fw_hat, void* f, hipStream_t s);
hipError_t launch_deconv_pad(const DeconvArgs& a, const void* f, void* fw_hat, hipStream_t s);

// ---- fft2d.hip: fused FFT + deconvolution of 2D and 3D FP32 plans on power-of-two grids
struct Fft2Plan {
  void* tw[3] = {nullptr, nullptr, nullptr};  // float2 twiddle tables, nf_d / 2 entries
  float2* mid = nullptr;   // 2D: n_trans * nm0 * nf1 (x-modes, transposed); 3D: first intermediate
  float2* mid2 = nullptr;  // 3D: second intermediate
  int64_t cap_mid = 0, cap_mid2 = 0;
  std::vector<void*> retired;
};
bool fft2_applicable(const DeconvArgs& a);
// twiddle tables and the intermediate buffer for a.n_trans transforms (no allocation at run time)
hipError_t fft2_plan_init(Fft2Plan& fp, const DeconvArgs& a);
void fft2_plan_free(Fft2Plan& fp);
// grid (FFT of the spread grid) -> deconvolved modes, and modes -> padded, deconvolved,
// transformed grid; isign selects the exponent sign (both unnormalised)
hipError_t fft2_shuffle(Fft2Plan& fp, const DeconvArgs& a, const void* grid, void* f, int isign, hipStream_t s);
hipError_t fft2_pad(Fft2Plan& fp, const DeconvArgs& a, const void* f, void* grid, int isign, hipStream_t s);

// ---- type 3 (elementwise.hip)
struct T3State {  // lives in device memory, filled by launch_t3_setup
  double C[3], D[3], gamma[3], h[3];
};
hipError_t launch_t3_setup(int dim, int prec, int64_t M, const void* const x[3], int64_t N, const void* const s[3],
                           const int nf[3], double upsampfac, T3State* st, double* scratch, int scratch_elems,
                           hipStream_t strm);
// x'_d = (x_d - C_d)/gamma_d (double), pre[i] = exp(i isign sum_d D_d x_d) (complex double)
hipError_t launch_t3_prep_src(int dim, int prec, int64_t M, const void* const x[3], const T3State* st, int isign,
                              double* xr[3], double* pre, hipStream_t s);
// s'_d = h_d gamma_d (s_d - D_d) (double), post[k] = exp(i isign sum (s-D) C) / prod phihat(s'_d)
struct QuadRule {
  int n;
  double z[40];
  double wphi[40];
};
hipError_t launch_t3_prep_tgt(int dim, int prec, int64_t N, const void* const s[3], const T3State* st, int isign,
                              const QuadRule& q, double* sr[3], double* post, hipStream_t strm);
// out[t][i] = in[t][i] * fac[i] (fac complex double), in/out complex T
hipError_t launch_cmul_pointwise(int prec, int64_t n_trans, int64_t M, const void* in, const double* fac, void* out,
                                 hipStream_t s);

}  // namespace nufftx
