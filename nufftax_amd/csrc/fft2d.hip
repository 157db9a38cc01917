// Fused uniform-grid stage of 2D FP32 type-1 and type-2 transforms on power-of-two fine
// grids (64 .. 4096 per axis): the c2c FFT of the fine grid (nufftax/transforms/nufft1.py:81,
// nufft2.py:74: fftn for isign < 0, unnormalised ifftn * N for isign > 0) together with the
// deconvolution, mode truncation and shift of deconvolve.py:174-224 (type 1) or the
// deconvolution and zero padding of deconvolve.py:227-304 (type 2).
//
// The 2D DFT is separable, and only the kept modes of one axis are needed before the other
// axis is transformed, so each direction is two passes over contiguous sequences:
//   type 1   rows:    grid[t][y][x]  -> mid[t][m0][y] = fac0[m0] * sum_x grid e(k0(m0) x / nf0)
//            columns: mid[t][m0][y] -> f[t][m1][m0]  = fac1[m1] * sum_y mid e(k1(m1) y / nf1)
//   type 2   columns: f[t][m1][m0]  -> mid[t][m0][y] = sum_m1 fac0 fac1 f e(k1(m1) y / nf1)
//            rows:    mid[t][m0][y] -> grid[t][y][x] = sum_m0 mid e(k0(m0) x / nf0)
// with e(u) = exp(isign 2 pi i u) and k(m) the FFT bin of mode index m (mode_to_grid).  The
// intermediate mid holds only nm0 of the nf0 x-bins, transposed so that both passes read and
// write whole sequences; a block handles R = 4 sequences (32-byte runs on the transposed side).
//
// Each sequence is transformed in LDS by an in-place radix-2 decimation-in-frequency FFT
// (results in bit-reversed order); twiddles exp(-2 pi i k / N), k < N/2, come from a table
// computed on the host in FP64 and are conjugated for isign > 0.
#include <cmath>
#include <vector>

#include "internal.hpp"

namespace nufftx {

namespace {

constexpr int kFftSeq = 4;         // sequences per block
constexpr int kFftThreads = 256;

__device__ __forceinline__ int fft_mode_to_bin(int m, int n, int nf, int modeord) {
  const int kmin = -(n / 2), kmax = (n - 1) / 2;
  const int k = modeord == 0 ? m + kmin : (m <= kmax ? m : m - n);
  return k >= 0 ? k : nf + k;
}

__device__ __forceinline__ int bitrev(int k, int logn) { return (int)(__brev((unsigned)k) >> (32 - logn)); }

// In-place DIF FFT of kFftSeq sequences of length N = 2^logn at buf[r * N + n]; twiddle table
// tw[k] = exp(-2 pi i k / N) in LDS.  Ends with a barrier.
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw, int logn, bool conj) {
  const int N = 1 << logn;
  const int nbf = kFftSeq * (N >> 1);
  for (int lh = logn - 1; lh >= 0; --lh) {
    const int half = 1 << lh;
    const int tstep = (N >> 1) >> lh;  // N / (2 half)
    for (int t = threadIdx.x; t < nbf; t += blockDim.x) {
      const int r = t >> (logn - 1), bf = t & ((N >> 1) - 1);
      const int pos = bf & (half - 1);
      const int i = r * N + ((bf >> lh) << (lh + 1)) + pos, j = i + half;
      const float2 a = buf[i], b = buf[j];
      float2 w = tw[pos * tstep];
      if (conj) w.y = -w.y;
      const float dr = a.x - b.x, di = a.y - b.y;
      buf[i] = make_float2(a.x + b.x, a.y + b.y);
      buf[j] = make_float2(dr * w.x - di * w.y, dr * w.y + di * w.x);
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void load_twiddles(float2* stw, const float2* __restrict__ tw, int n) {
  for (int k = threadIdx.x; k < (n >> 1); k += blockDim.x) stw[k] = tw[k];
}

struct Fft2K {
  int logn[2];     // log2 nf0, log2 nf1
  int nf[2], nm[2];
  int modeord;
  const float* fac[2];   // deconvolution factors in output mode order
  const float2* tw[2];   // twiddle tables, nf_d / 2 entries
  int64_t grid_cells, mode_cells, mid_cells;  // per transform
};

// type 1, rows: grid rows y0 .. y0+R-1 -> mid[m0][y] (kept x-modes, times fac0)
__global__ void __launch_bounds__(kFftThreads) k_fft2_rows_t1(Fft2K k, const float2* __restrict__ grid,
                                                              float2* __restrict__ mid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  const int N = k.nf[0], ny = k.nf[1];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int y0 = blockIdx.x * kFftSeq;
  const int64_t t = blockIdx.y;
  const float2* g = grid + t * k.grid_cells;
  load_twiddles(stw, k.tw[0], N);
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) {
    const int r = e / N, x = e - r * N;
    buf[e] = y0 + r < ny ? g[(int64_t)(y0 + r) * N + x] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft(buf, stw, k.logn[0], conj != 0);
  float2* o = mid + t * k.mid_cells;
  for (int e = threadIdx.x; e < kFftSeq * k.nm[0]; e += blockDim.x) {
    const int m0 = e / kFftSeq, r = e - m0 * kFftSeq;
    if (y0 + r >= ny) continue;
    const float2 v = buf[r * N + bitrev(fft_mode_to_bin(m0, k.nm[0], N, k.modeord), k.logn[0])];
    const float f = k.fac[0][m0];
    o[(int64_t)m0 * ny + y0 + r] = make_float2(v.x * f, v.y * f);
  }
}

// type 1, columns: mid[m0 .. m0+R-1][y] -> f[m1][m0] (kept y-modes, times fac1)
__global__ void __launch_bounds__(kFftThreads) k_fft2_cols_t1(Fft2K k, const float2* __restrict__ mid,
                                                              float2* __restrict__ f, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  const int N = k.nf[1], nm0 = k.nm[0];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int c0 = blockIdx.x * kFftSeq;
  const int64_t t = blockIdx.y;
  const float2* in = mid + t * k.mid_cells;
  load_twiddles(stw, k.tw[1], N);
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) {
    const int r = e / N, y = e - r * N;
    buf[e] = c0 + r < nm0 ? in[(int64_t)(c0 + r) * N + y] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft(buf, stw, k.logn[1], conj != 0);
  float2* o = f + t * k.mode_cells;
  for (int e = threadIdx.x; e < kFftSeq * k.nm[1]; e += blockDim.x) {
    const int m1 = e / kFftSeq, r = e - m1 * kFftSeq;
    if (c0 + r >= nm0) continue;
    const float2 v = buf[r * N + bitrev(fft_mode_to_bin(m1, k.nm[1], N, k.modeord), k.logn[1])];
    const float fa = k.fac[1][m1];
    o[(int64_t)m1 * nm0 + c0 + r] = make_float2(v.x * fa, v.y * fa);
  }
}

// type 2, columns: f[m1][m0 .. m0+R-1] placed at the y-bins (times fac0 fac1) -> mid[m0][y]
__global__ void __launch_bounds__(kFftThreads) k_fft2_cols_t2(Fft2K k, const float2* __restrict__ f,
                                                              float2* __restrict__ mid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  const int N = k.nf[1], nm0 = k.nm[0];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int c0 = blockIdx.x * kFftSeq;
  const int64_t t = blockIdx.y;
  const float2* in = f + t * k.mode_cells;
  load_twiddles(stw, k.tw[1], N);
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int e = threadIdx.x; e < kFftSeq * k.nm[1]; e += blockDim.x) {
    const int m1 = e / kFftSeq, r = e - m1 * kFftSeq;
    if (c0 + r >= nm0) continue;
    const float2 v = in[(int64_t)m1 * nm0 + c0 + r];
    const float fa = k.fac[0][c0 + r] * k.fac[1][m1];
    buf[r * N + fft_mode_to_bin(m1, k.nm[1], N, k.modeord)] = make_float2(v.x * fa, v.y * fa);
  }
  __syncthreads();
  lds_fft(buf, stw, k.logn[1], conj != 0);
  float2* o = mid + t * k.mid_cells;
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) {
    const int r = e / N, y = e - r * N;
    if (c0 + r < nm0) o[(int64_t)(c0 + r) * N + y] = buf[r * N + bitrev(y, k.logn[1])];
  }
}

// type 2, rows: mid[m0][y0 .. y0+R-1] placed at the x-bins -> grid rows y0 .. y0+R-1
__global__ void __launch_bounds__(kFftThreads) k_fft2_rows_t2(Fft2K k, const float2* __restrict__ mid,
                                                              float2* __restrict__ grid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  const int N = k.nf[0], ny = k.nf[1];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int y0 = blockIdx.x * kFftSeq;
  const int64_t t = blockIdx.y;
  const float2* in = mid + t * k.mid_cells;
  load_twiddles(stw, k.tw[0], N);
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  for (int e = threadIdx.x; e < kFftSeq * k.nm[0]; e += blockDim.x) {
    const int m0 = e / kFftSeq, r = e - m0 * kFftSeq;
    if (y0 + r >= ny) continue;
    buf[r * N + fft_mode_to_bin(m0, k.nm[0], N, k.modeord)] = in[(int64_t)m0 * ny + y0 + r];
  }
  __syncthreads();
  lds_fft(buf, stw, k.logn[0], conj != 0);
  float2* g = grid + t * k.grid_cells;
  for (int e = threadIdx.x; e < kFftSeq * N; e += blockDim.x) {
    const int r = e / N, x = e - r * N;
    if (y0 + r < ny) g[(int64_t)(y0 + r) * N + x] = buf[r * N + bitrev(x, k.logn[0])];
  }
}

int log2_exact(int n) {
  int l = 0;
  while ((1 << l) < n) ++l;
  return (1 << l) == n ? l : -1;
}

Fft2K make_k(const Fft2Plan& fp, const DeconvArgs& a) {
  Fft2K k;
  for (int d = 0; d < 2; ++d) {
    k.nf[d] = a.nf[d];
    k.nm[d] = a.nm[d];
    k.logn[d] = log2_exact(a.nf[d]);
    k.fac[d] = static_cast<const float*>(a.fac[d]);
    k.tw[d] = static_cast<const float2*>(fp.tw[d]);
  }
  k.modeord = a.modeord;
  k.grid_cells = (int64_t)a.nf[0] * a.nf[1];
  k.mode_cells = (int64_t)a.nm[0] * a.nm[1];
  k.mid_cells = (int64_t)a.nm[0] * a.nf[1];
  return k;
}

size_t fft_lds(int n) { return (size_t)(n / 2 + kFftSeq * n) * sizeof(float2); }

// LDS of a pass over length-n sequences; grids of 2048 and 4096 need more than 64 KB
void fft_lds_attr(const void* kern, int n) {
  const size_t b = fft_lds(n);
  if (b > 65536) (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
}

}  // namespace

bool fft2_applicable(const DeconvArgs& a) {
  if (a.dim != 2 || a.prec != kF32 || a.from_phihat) return false;
  for (int d = 0; d < 2; ++d) {
    const int l = log2_exact(a.nf[d]);
    if (l < 6 || l > 12 || a.nm[d] > a.nf[d]) return false;
  }
  return true;
}

hipError_t fft2_plan_init(Fft2Plan& fp, const int* nf) {
  for (int d = 0; d < 2; ++d) {
    const int n = nf[d];
    std::vector<float2> h((size_t)n / 2);
    for (int k = 0; k < n / 2; ++k) {
      const double ang = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
      h[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2* p = nullptr;
    hipError_t e = hipMalloc((void**)&p, h.size() * sizeof(float2));
    if (e != hipSuccess) return e;
    if ((e = hipMemcpy(p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
      (void)hipFree(p);
      return e;
    }
    fp.tw[d] = p;
  }
  return hipSuccess;
}

void fft2_plan_free(Fft2Plan& fp) {
  for (int d = 0; d < 2; ++d) {
    if (fp.tw[d]) (void)hipFree(fp.tw[d]);
    fp.tw[d] = nullptr;
  }
  for (void* p : fp.retired) (void)hipFree(p);
  fp.retired.clear();
  if (fp.mid) (void)hipFree(fp.mid);
  fp.mid = nullptr;
  fp.cap_mid = 0;
}

hipError_t fft2_shuffle(Fft2Plan& fp, const DeconvArgs& a, const void* grid, void* f, int isign, hipStream_t s) {
  const Fft2K k = make_k(fp, a);
  hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * k.mid_cells, fp.retired);
  if (e != hipSuccess) return e;
  const int conj = isign > 0 ? 1 : 0;
  const unsigned nt = (unsigned)a.n_trans;
  fft_lds_attr((const void*)k_fft2_rows_t1, a.nf[0]);
  fft_lds_attr((const void*)k_fft2_cols_t1, a.nf[1]);
  hipLaunchKernelGGL(k_fft2_rows_t1, dim3((a.nf[1] + kFftSeq - 1) / kFftSeq, nt), dim3(kFftThreads), fft_lds(a.nf[0]),
                     s, k, static_cast<const float2*>(grid), fp.mid, conj);
  hipLaunchKernelGGL(k_fft2_cols_t1, dim3((a.nm[0] + kFftSeq - 1) / kFftSeq, nt), dim3(kFftThreads), fft_lds(a.nf[1]),
                     s, k, fp.mid, static_cast<float2*>(f), conj);
  return hipGetLastError();
}

hipError_t fft2_pad(Fft2Plan& fp, const DeconvArgs& a, const void* f, void* grid, int isign, hipStream_t s) {
  const Fft2K k = make_k(fp, a);
  hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * k.mid_cells, fp.retired);
  if (e != hipSuccess) return e;
  const int conj = isign > 0 ? 1 : 0;
  const unsigned nt = (unsigned)a.n_trans;
  fft_lds_attr((const void*)k_fft2_cols_t2, a.nf[1]);
  fft_lds_attr((const void*)k_fft2_rows_t2, a.nf[0]);
  hipLaunchKernelGGL(k_fft2_cols_t2, dim3((a.nm[0] + kFftSeq - 1) / kFftSeq, nt), dim3(kFftThreads), fft_lds(a.nf[1]),
                     s, k, static_cast<const float2*>(f), fp.mid, conj);
  hipLaunchKernelGGL(k_fft2_rows_t2, dim3((a.nf[1] + kFftSeq - 1) / kFftSeq, nt), dim3(kFftThreads), fft_lds(a.nf[0]),
                     s, k, fp.mid, static_cast<float2*>(grid), conj);
  return hipGetLastError();
}

}  // namespace nufftx
