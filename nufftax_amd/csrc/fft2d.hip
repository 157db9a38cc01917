// Fused uniform-grid stage of 1D, 2D and 3D FP32 type-1 and type-2 transforms on power-of-two
// fine grids (64 .. 4096 per axis; 3D: k_fft3_pass, three passes, described there; 1D: one
// k_fft3_pass per transform): the c2c FFT of
// the fine grid (nufftax/transforms/nufft1.py:81, nufft2.py:74: fftn for isign < 0,
// unnormalised ifftn * N for isign > 0) together with the
// deconvolution, mode truncation and shift of deconvolve.py:174-224 (type 1) or the
// deconvolution and zero padding of deconvolve.py:227-304 (type 2).
//
// The 2D DFT is separable, and only the kept modes of one axis are needed before the other
// axis is transformed, so each direction is two passes over contiguous sequences:
//   type 1   rows:    grid[t][y][x]  -> mid[t][m0][y] = fac0[m0] * sum_x grid e(k0(m0) x / nf0)
//            columns: mid[t][m0][y] -> f[t][m1][m0]  = fac1[m1] * sum_y mid e(k1(m1) y / nf1)
//   type 2   columns: f[t][m1][m0]  -> mid[t][m0][y] = sum_m1 fac0 fac1 f e(k1(m1) y / nf1)
//            rows:    mid[t][m0][y] -> grid[t][y][x] = sum_m0 mid e(k0(m0) x / nf0)
// with e(u) = exp(isign 2 pi i u) and k(m) the FFT bin of mode index m.  The intermediate mid
// holds only nm0 of the nf0 x-bins, transposed so that both passes read and write whole
// sequences; a block handles S sequences (S 8-byte runs on the transposed side).
//
// Each sequence of N points is transformed in LDS by the Stockham autosort FFT (Govindaraju
// et al., "High performance discrete Fourier transforms on graphics processors", SC 2008) in
// 2 or 3 passes of radix 8 or 16 (1024 = 16 x 8 x 8): pass p reads x[j + r N/R] (r < R), scales
// them by W_(Ns R)^(r (j mod Ns)), takes an R-point register DFT and writes it to
// (j / Ns) Ns R + (j mod Ns) + r Ns, Ns being the product of the earlier radices; the output of
// the last pass is in natural order.  A register DFT is a radix-2 decimation-in-time FFT on
// bit-reversed loads.  Twiddles exp(-2 pi i k / N), k < N/2, come from a table computed on the
// host in FP64 (staged in LDS) and are conjugated for isign > 0.  A block of kFftThreads
// threads transforms S sequences with T = min(N / 8, 128) threads each (one 8-point column of
// a radix-8 pass per thread; the 3D passes take as many threads as the pass with the fewest
// columns has columns, 16 for N = 256).  An earlier four-step version (two passes of 32-point register
// DFTs, 32 threads per 1024-point sequence) left most of the GPU idle: 128 blocks of two waves
// for the 512 kept columns of a 1024^2 grid.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "internal.hpp"

namespace nufftx {

namespace {

constexpr int kFftThreads = 256;

// threads per sequence of 2^logn points, and sequences per block
__host__ __device__ constexpr int fft_seq_threads(int logn) {
  return (1 << logn) / 8 < 8 ? 8 : ((1 << logn) / 8 > 128 ? 128 : (1 << logn) / 8);
}
__host__ __device__ constexpr int fft_seqs(int logn) { return kFftThreads / fft_seq_threads(logn); }

// 3D passes: as many threads per sequence as the Stockham pass with the fewest columns has
// columns (every thread busy in every pass; 16 for 256 points, against 32 above), 8 .. 128
__host__ __device__ constexpr int fft3_seq_threads(int logn) {
  return ((1 << logn) >> (logn / ((logn + 3) / 4))) < 8
             ? 8
             : (((1 << logn) >> (logn / ((logn + 3) / 4))) > 128 ? 128 : ((1 << logn) >> (logn / ((logn + 3) / 4))));
}
__host__ __device__ constexpr int fft3_seqs(int logn) { return kFftThreads / fft3_seq_threads(logn); }

__device__ __forceinline__ int fft_mode_to_bin(int m, int n, int nf, int modeord) {
  const int kmin = -(n / 2), kmax = (n - 1) / 2;
  const int k = modeord == 0 ? m + kmin : (m <= kmax ? m : m - n);
  return k >= 0 ? k : nf + k;
}

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// W_N^m for 0 <= m < N from the half table tw[k] = exp(-2 pi i k / N), conjugated if conj
__device__ __forceinline__ float2 twiddle(const float2* tw, int m, int N, bool conj) {
  const int h = N >> 1;
  float2 w = m < h ? tw[m] : make_float2(-tw[m - h].x, -tw[m - h].y);
  if (conj) w.y = -w.y;
  return w;
}

template <int L>
__host__ __device__ constexpr int bitrev_c(int i) {
  int r = 0;
  for (int b = 0; b < L; ++b) r |= ((i >> b) & 1) << (L - 1 - b);
  return r;
}

// LDS layout of a sequence of N = 2^LOGN points: one pad slot after every 2^ceil(LOGN / 2)
// points, which spreads the strided writes of the first Stockham passes over the banks; EXTRA
// more slots at the end of each sequence (1 makes the stride odd, so that reads across
// sequences at one index fall in different banks)
template <int LOGN, int EXTRA = 0>
struct FftLayout {
  static constexpr int L1 = (LOGN + 1) / 2;
  static constexpr int N = 1 << LOGN;
  static constexpr int PL = N + (N >> L1) + EXTRA;  // padded sequence length
  static __device__ __forceinline__ int at(int r, int n) { return r * PL + n + (n >> L1); }
};

// In-register DFT of R = 2^L points held in bit-reversed order (v[bitrev(n)] = x[n]); result
// in natural order.  W_R^j = W_N^(j N / R).
template <int L>
__device__ __forceinline__ void reg_dft(float2 (&v)[1 << L], const float2* tw, int N, bool conj) {
  constexpr int R = 1 << L;
  const int step = N / R;
#pragma unroll
  for (int len = 2; len <= R; len <<= 1) {
    const int half = len >> 1;
#pragma unroll
    for (int j = 0; j < half; ++j) {
      const float2 w = twiddle(tw, j * (R / len) * step, N, conj);
#pragma unroll
      for (int i = 0; i < R; i += len) {
        const float2 a = v[i + j], b = cmul(v[i + j + half], w);
        v[i + j] = make_float2(a.x + b.x, a.y + b.y);
        v[i + j + half] = make_float2(a.x - b.x, a.y - b.y);
      }
    }
  }
}

// One Stockham pass of radix R = 2^L over the S sequences of a block (in place in the padded
// layout: every read of the pass precedes every write); Ns = 2^LNS.  Ends with a barrier.
template <int LOGN, int L, int LNS, int T = fft_seq_threads(LOGN), class Lay = FftLayout<LOGN>>
__device__ __forceinline__ void stockham_pass(float2* buf, const float2* tw, bool conj) {
  constexpr int N = 1 << LOGN, R = 1 << L, NR = N / R, Ns = 1 << LNS;
  constexpr int J = (NR + T - 1) / T;  // columns per thread
  const int seq = threadIdx.x / T, q = threadIdx.x % T;
  float2 v[J][R];
#pragma unroll
  for (int jj = 0; jj < J; ++jj) {
    const int j = q + jj * T;
    if (j >= NR) break;
    const int k = j & (Ns - 1);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float2 x = buf[Lay::at(seq, j + r * NR)];
      v[jj][bitrev_c<L>(r)] = (r == 0 || LNS == 0) ? x : cmul(x, twiddle(tw, r * k * (N / (Ns * R)), N, conj));
    }
    reg_dft<L>(v[jj], tw, N, conj);
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < J; ++jj) {
    const int j = q + jj * T;
    if (j >= NR) break;
    const int k = j & (Ns - 1);
    const int d = (j >> LNS) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) buf[Lay::at(seq, d + r * Ns)] = v[jj][r];
  }
  __syncthreads();
}

// In-place FFT of the S sequences of N = 2^LOGN points (6 <= LOGN <= 12) of a block, natural
// order in and out: ceil(LOGN / 4) passes of 3 or 4 bits (the first ones the larger).  Ends
// with a barrier.  T threads per sequence, layout Lay.
template <int LOGN, int T = fft_seq_threads(LOGN), class Lay = FftLayout<LOGN>>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw, bool conj) {
  constexpr int NP = (LOGN + 3) / 4;
  constexpr int B0 = LOGN / NP + (0 < LOGN % NP ? 1 : 0);
  constexpr int B1 = LOGN / NP + (1 < LOGN % NP ? 1 : 0);
  constexpr int B2 = LOGN - B0 - B1;
  static_assert(NP >= 2 && NP <= 3 && B0 <= 4 && B1 <= 4 && B2 <= 4, "2 or 3 passes of radix <= 16");
  stockham_pass<LOGN, B0, 0, T, Lay>(buf, tw, conj);
  stockham_pass<LOGN, B1, B0, T, Lay>(buf, tw, conj);
  if constexpr (B2 > 0) stockham_pass<LOGN, B2, B0 + B1, T, Lay>(buf, tw, conj);
}

__device__ __forceinline__ void load_twiddles(float2* stw, const float2* __restrict__ tw, int n) {
  for (int k = threadIdx.x; k < (n >> 1); k += kFftThreads) stw[k] = tw[k];
}

struct Fft2K {
  int logn[2];     // log2 nf0, log2 nf1
  int nf[2], nm[2];
  int modeord;
  const float* fac[2];   // deconvolution factors in output mode order
  const float2* tw[2];   // twiddle tables, nf_d / 2 entries
  int64_t grid_cells, mode_cells, mid_cells;  // per transform
};

// type 1, rows: grid rows y0 .. y0+S-1 -> mid[m0][y] (kept x-modes, times fac0)
template <int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft2_rows_t1(Fft2K k, const float2* __restrict__ grid,
                                                              float2* __restrict__ mid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  constexpr int N = 1 << LOGN, S = fft_seqs(LOGN);
  const int ny = k.nf[1];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int y0 = blockIdx.x * S;
  const int64_t t = blockIdx.y;
  const float2* g = grid + t * k.grid_cells;
  load_twiddles(stw, k.tw[0], N);
#pragma unroll
  for (int u = 0; u < S * N / kFftThreads; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    const int r = e / N, x = e - r * N;
    buf[FftLayout<LOGN>::at(r, x)] = y0 + r < ny ? g[(int64_t)(y0 + r) * N + x] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<LOGN>(buf, stw, conj != 0);
  float2* o = mid + t * k.mid_cells;
#pragma unroll 4
  for (int e = threadIdx.x; e < S * k.nm[0]; e += kFftThreads) {
    const int m0 = e / S, r = e - m0 * S;
    if (y0 + r >= ny) continue;
    const float2 v = buf[FftLayout<LOGN>::at(r, fft_mode_to_bin(m0, k.nm[0], N, k.modeord))];
    const float f = k.fac[0][m0];
    o[(int64_t)m0 * ny + y0 + r] = make_float2(v.x * f, v.y * f);
  }
}

// type 1, columns: mid[m0 .. m0+S-1][y] -> f[m1][m0] (kept y-modes, times fac1)
template <int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft2_cols_t1(Fft2K k, const float2* __restrict__ mid,
                                                              float2* __restrict__ f, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  constexpr int N = 1 << LOGN, S = fft_seqs(LOGN);
  const int nm0 = k.nm[0];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int c0 = blockIdx.x * S;
  const int64_t t = blockIdx.y;
  const float2* in = mid + t * k.mid_cells;
  load_twiddles(stw, k.tw[1], N);
#pragma unroll
  for (int u = 0; u < S * N / kFftThreads; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    const int r = e / N, y = e - r * N;
    buf[FftLayout<LOGN>::at(r, y)] = c0 + r < nm0 ? in[(int64_t)(c0 + r) * N + y] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  lds_fft<LOGN>(buf, stw, conj != 0);
  float2* o = f + t * k.mode_cells;
#pragma unroll 4
  for (int e = threadIdx.x; e < S * k.nm[1]; e += kFftThreads) {
    const int m1 = e / S, r = e - m1 * S;
    if (c0 + r >= nm0) continue;
    const float2 v = buf[FftLayout<LOGN>::at(r, fft_mode_to_bin(m1, k.nm[1], N, k.modeord))];
    const float fa = k.fac[1][m1];
    o[(int64_t)m1 * nm0 + c0 + r] = make_float2(v.x * fa, v.y * fa);
  }
}

// type 2, columns: f[m1][m0 .. m0+S-1] placed at the y-bins (times fac0 fac1) -> mid[m0][y]
template <int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft2_cols_t2(Fft2K k, const float2* __restrict__ f,
                                                              float2* __restrict__ mid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  constexpr int N = 1 << LOGN, S = fft_seqs(LOGN);
  const int nm0 = k.nm[0];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int c0 = blockIdx.x * S;
  const int64_t t = blockIdx.y;
  const float2* in = f + t * k.mode_cells;
  load_twiddles(stw, k.tw[1], N);
#pragma unroll
  for (int u = 0; u < S * FftLayout<LOGN>::PL / kFftThreads + 1; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    if (e < S * FftLayout<LOGN>::PL) buf[e] = make_float2(0.f, 0.f);
  }
  __syncthreads();
#pragma unroll 4
  for (int e = threadIdx.x; e < S * k.nm[1]; e += kFftThreads) {
    const int m1 = e / S, r = e - m1 * S;
    if (c0 + r >= nm0) continue;
    const float2 v = in[(int64_t)m1 * nm0 + c0 + r];
    const float fa = k.fac[0][c0 + r] * k.fac[1][m1];
    buf[FftLayout<LOGN>::at(r, fft_mode_to_bin(m1, k.nm[1], N, k.modeord))] = make_float2(v.x * fa, v.y * fa);
  }
  __syncthreads();
  lds_fft<LOGN>(buf, stw, conj != 0);
  float2* o = mid + t * k.mid_cells;
#pragma unroll
  for (int u = 0; u < S * N / kFftThreads; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    const int r = e / N, y = e - r * N;
    if (c0 + r < nm0) o[(int64_t)(c0 + r) * N + y] = buf[FftLayout<LOGN>::at(r, y)];
  }
}

// type 2, rows: mid[m0][y0 .. y0+S-1] placed at the x-bins -> grid rows y0 .. y0+S-1
template <int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft2_rows_t2(Fft2K k, const float2* __restrict__ mid,
                                                              float2* __restrict__ grid, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  constexpr int N = 1 << LOGN, S = fft_seqs(LOGN);
  const int ny = k.nf[1];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int y0 = blockIdx.x * S;
  const int64_t t = blockIdx.y;
  const float2* in = mid + t * k.mid_cells;
  load_twiddles(stw, k.tw[0], N);
#pragma unroll
  for (int u = 0; u < S * FftLayout<LOGN>::PL / kFftThreads + 1; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    if (e < S * FftLayout<LOGN>::PL) buf[e] = make_float2(0.f, 0.f);
  }
  __syncthreads();
#pragma unroll 4
  for (int e = threadIdx.x; e < S * k.nm[0]; e += kFftThreads) {
    const int m0 = e / S, r = e - m0 * S;
    if (y0 + r >= ny) continue;
    buf[FftLayout<LOGN>::at(r, fft_mode_to_bin(m0, k.nm[0], N, k.modeord))] = in[(int64_t)m0 * ny + y0 + r];
  }
  __syncthreads();
  lds_fft<LOGN>(buf, stw, conj != 0);
  float2* g = grid + t * k.grid_cells;
#pragma unroll
  for (int u = 0; u < S * N / kFftThreads; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    const int r = e / N, x = e - r * N;
    if (y0 + r < ny) g[(int64_t)(y0 + r) * N + x] = buf[FftLayout<LOGN>::at(r, x)];
  }
}

// One pass of the 3D fused FFT: S sequences per block along one axis, read contiguously from
// `in` and written with stride `ostride` to `out` (consecutive sequences to consecutive
// addresses, so a block writes runs of S values).  Sequence seq = a D1 + b starts at
// a P1 + b P2 in `in` and a Q1 + b Q2 in `out`.  Inputs are either N grid values or the nin
// modes of this axis (placed at their FFT bins, the rest zero); outputs are the N bins or the
// nout kept modes of this axis.  fac: facA[mode of this axis] facB[b] facC[a] (each optional)
// on the modes side.
struct Fft3Pass {
  int nin, nout, nmode, modeord;
  int in_modes, out_modes;
  int64_t nseq, D1, P1, P2, Q1, Q2, ostride, in_t, out_t;
  const float* facA;
  const float* facB;
  const float* facC;
  const float2* tw;
};

template <int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft3_pass(Fft3Pass P, const float2* __restrict__ in,
                                                           float2* __restrict__ out, int conj) {
  extern __shared__ __align__(16) float2 smem[];
  constexpr int N = 1 << LOGN, T = fft3_seq_threads(LOGN), S = kFftThreads / T;
  constexpr int EPT = S * N / kFftThreads;  // values per thread in and out
  using Lay = FftLayout<LOGN, 1>;
  __shared__ int64_t sa[S], sb[S];
  float2* stw = smem;
  float2* buf = smem + (N >> 1);
  const int64_t s0 = (int64_t)blockIdx.x * S;
  const int64_t t = blockIdx.y;
  const float2* inT = in + t * P.in_t;
  float2* outT = out + t * P.out_t;
  if ((int)threadIdx.x < S) {
    const int64_t seq = s0 + threadIdx.x;
    sa[threadIdx.x] = seq / P.D1;
    sb[threadIdx.x] = seq % P.D1;
  }
  load_twiddles(stw, P.tw, N);
  if (P.in_modes)
    for (int e = threadIdx.x; e < S * Lay::PL; e += kFftThreads) buf[e] = make_float2(0.f, 0.f);
  __syncthreads();
  // every load of the thread is issued before the first LDS write
  float2 v[EPT];
  if (P.in_modes) {
    const int nin = P.nin;
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = threadIdx.x + u * kFftThreads;
      const int r = e / nin, i = e - r * nin;
      v[u] = make_float2(0.f, 0.f);
      if (r < S && s0 + r < P.nseq) v[u] = inT[sa[r] * P.P1 + sb[r] * P.P2 + i];
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = threadIdx.x + u * kFftThreads;
      const int r = e / nin, i = e - r * nin;
      if (r < S && s0 + r < P.nseq) {
        float f = P.facA ? P.facA[i] : 1.0f;
        if (P.facB) f *= P.facB[sb[r]];
        if (P.facC) f *= P.facC[sa[r]];
        buf[Lay::at(r, fft_mode_to_bin(i, P.nmode, N, P.modeord))] = make_float2(v[u].x * f, v[u].y * f);
      }
    }
  } else {
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = threadIdx.x + u * kFftThreads;
      const int r = e >> LOGN, i = e & (N - 1);
      v[u] = make_float2(0.f, 0.f);
      if (s0 + r < P.nseq) v[u] = inT[sa[r] * P.P1 + sb[r] * P.P2 + i];
    }
#pragma unroll
    for (int u = 0; u < EPT; ++u) {
      const int e = threadIdx.x + u * kFftThreads;
      buf[Lay::at(e >> LOGN, e & (N - 1))] = v[u];
    }
  }
  __syncthreads();
  lds_fft<LOGN, T, Lay>(buf, stw, conj != 0);
  // consecutive lanes take consecutive sequences at one output index: runs of S values
#pragma unroll
  for (int u = 0; u < EPT; ++u) {
    const int e = threadIdx.x + u * kFftThreads;
    const int j = e / S, r = e & (S - 1);
    if (j >= P.nout || s0 + r >= P.nseq) continue;
    float2 w;
    if (P.out_modes) {
      w = buf[Lay::at(r, fft_mode_to_bin(j, P.nmode, N, P.modeord))];
      float f = P.facA ? P.facA[j] : 1.0f;
      if (P.facB) f *= P.facB[sb[r]];
      if (P.facC) f *= P.facC[sa[r]];
      w = make_float2(w.x * f, w.y * f);
    } else {
      w = buf[Lay::at(r, j)];
    }
    outT[sa[r] * P.Q1 + sb[r] * P.Q2 + (int64_t)j * P.ostride] = w;
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

// LDS of a pass over length-n sequences (twiddles + S sequences)
size_t fft_lds(int n) {
  const int l = log2_exact(n);
  const int pl = n + (n >> ((l + 1) / 2));
  return (size_t)(n / 2 + fft_seqs(l) * pl) * sizeof(float2);
}

// LDS of a 3D pass over length-n sequences (twiddles + S sequences of odd stride)
size_t fft3_lds(int n) {
  const int l = log2_exact(n);
  const int pl = n + (n >> ((l + 1) / 2)) + 1;
  return (size_t)(n / 2 + fft3_seqs(l) * pl) * sizeof(float2);
}

unsigned fft_blocks(int count, int n) {
  const int S = fft_seqs(log2_exact(n));
  return (unsigned)((count + S - 1) / S);
}

void fft_lds_attr(const void* kern, size_t b) {
  if (b > 65536) (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b);
}

}  // namespace

bool fft2_applicable(const DeconvArgs& a) {
  if (a.dim < 1 || a.dim > 3 || a.prec != kF32 || a.from_phihat) return false;
  if (a.dim == 3)
    if (const char* e = std::getenv("NUFFTX_FFT3_FUSED"))
      if (e[0] == '0') return false;
  if (a.dim == 1) {
    // one block per transform with a single busy sequence: it removes a launch and rocFFT's
    // host-side cost per call (1D 100k points, 1024 modes: 57 -> 50 us through the public
    // transform, +3 us of GPU time); large batches of short transforms stay with rocFFT
    if (a.n_trans > 64) return false;
    if (const char* e = std::getenv("NUFFTX_FFT1_FUSED"))
      if (e[0] == '0') return false;
  }
  for (int d = 0; d < a.dim; ++d) {
    const int l = log2_exact(a.nf[d]);
    if (l < 6 || l > 12 || a.nm[d] > a.nf[d]) return false;
  }
  return true;
}

// intermediate sizes of the 3D passes per transform: type 2 writes [n2][nf0][n1] then
// [nf1][nf0][n2], type 1 [nf2][n0][nf1] then [n1][n0][nf2]
int64_t fft3_mid_cells(const DeconvArgs& a) {
  const int64_t n0 = a.nm[0], n1 = a.nm[1], n2 = a.nm[2], f0 = a.nf[0], f1 = a.nf[1], f2 = a.nf[2];
  return std::max(n2 * f0 * n1, f2 * n0 * f1);
}
int64_t fft3_mid2_cells(const DeconvArgs& a) {
  const int64_t n0 = a.nm[0], n1 = a.nm[1], n2 = a.nm[2], f0 = a.nf[0], f1 = a.nf[1], f2 = a.nf[2];
  return std::max(f1 * f0 * n2, n1 * n0 * f2);
}

hipError_t fft2_plan_init(Fft2Plan& fp, const DeconvArgs& a) {
  hipError_t e;
  if (a.dim == 3) {
    if ((e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * fft3_mid_cells(a), fp.retired)) != hipSuccess) return e;
    if ((e = grow_buffer(fp.mid2, fp.cap_mid2, a.n_trans * fft3_mid2_cells(a), fp.retired)) != hipSuccess) return e;
  } else if (a.dim == 2 &&
             (e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * (int64_t)a.nm[0] * a.nf[1], fp.retired)) != hipSuccess) {
    return e;
  }
  for (int d = 0; d < a.dim; ++d) {
    const int n = a.nf[d];
    std::vector<float2> h((size_t)n / 2);
    for (int k = 0; k < n / 2; ++k) {
      const double ang = -2.0 * 3.14159265358979323846 * (double)k / (double)n;
      h[k] = make_float2((float)std::cos(ang), (float)std::sin(ang));
    }
    float2* p = nullptr;
    if ((e = dev_malloc((void**)&p, h.size() * sizeof(float2))) != hipSuccess) return e;
    if ((e = hipMemcpy(p, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice)) != hipSuccess) {
      (void)hipFree(p);
      return e;
    }
    fp.tw[d] = p;
  }
  return hipSuccess;
}

void fft2_plan_free(Fft2Plan& fp) {
  for (int d = 0; d < 3; ++d) {
    if (fp.tw[d]) (void)hipFree(fp.tw[d]);
    fp.tw[d] = nullptr;
  }
  for (void* p : fp.retired) (void)hipFree(p);
  fp.retired.clear();
  if (fp.mid) (void)hipFree(fp.mid);
  if (fp.mid2) (void)hipFree(fp.mid2);
  fp.mid = nullptr;
  fp.mid2 = nullptr;
  fp.cap_mid = fp.cap_mid2 = 0;
}

// launch kern<LOGN> for the runtime log2 of the transformed axis
template <template <int> class K, typename... Args>
void launch_logn(int logn, dim3 grid, size_t lds, hipStream_t s, Args... args) {
  switch (logn) {
    case 6: K<6>::launch(grid, lds, s, args...); break;
    case 7: K<7>::launch(grid, lds, s, args...); break;
    case 8: K<8>::launch(grid, lds, s, args...); break;
    case 9: K<9>::launch(grid, lds, s, args...); break;
    case 10: K<10>::launch(grid, lds, s, args...); break;
    case 11: K<11>::launch(grid, lds, s, args...); break;
    default: K<12>::launch(grid, lds, s, args...); break;
  }
}

#define FFT2_LAUNCHER(name)                                                                    \
  template <int L>                                                                             \
  struct name##_l {                                                                            \
    template <typename... Args>                                                                \
    static void launch(dim3 grid, size_t lds, hipStream_t s, Args... args) {                   \
      fft_lds_attr((const void*)name<L>, lds);                                                 \
      hipLaunchKernelGGL(name<L>, grid, dim3(kFftThreads), lds, s, args...);                   \
    }                                                                                          \
  };
FFT2_LAUNCHER(k_fft2_rows_t1)
FFT2_LAUNCHER(k_fft2_cols_t1)
FFT2_LAUNCHER(k_fft2_cols_t2)
FFT2_LAUNCHER(k_fft2_rows_t2)
FFT2_LAUNCHER(k_fft3_pass)
#undef FFT2_LAUNCHER

// 3D: three passes (launch one per axis; logn of the transformed axis)
hipError_t fft3_run(Fft2Plan& fp, const DeconvArgs& a, const void* src, void* dst, int isign, bool to_grid,
                    hipStream_t s) {
  const int64_t n0 = a.nm[0], n1 = a.nm[1], n2 = a.nm[2], f0 = a.nf[0], f1 = a.nf[1], f2 = a.nf[2];
  const int conj = isign > 0 ? 1 : 0;
  const unsigned nt = (unsigned)a.n_trans;
  const int64_t mid_t = fft3_mid_cells(a), mid2_t = fft3_mid2_cells(a);
  const int64_t grid_t = f0 * f1 * f2, mode_t = n0 * n1 * n2;
  const float* fac[3] = {static_cast<const float*>(a.fac[0]), static_cast<const float*>(a.fac[1]),
                         static_cast<const float*>(a.fac[2])};
  auto pass = [&](int axis, Fft3Pass P, const float2* in, float2* out) {
    const int logn = log2_exact(a.nf[axis]);
    P.modeord = a.modeord;
    P.nmode = a.nm[axis];
    P.tw = static_cast<const float2*>(fp.tw[axis]);
    const unsigned blocks = (unsigned)((P.nseq + fft3_seqs(logn) - 1) / fft3_seqs(logn));
    launch_logn<k_fft3_pass_l>(logn, dim3(blocks, nt), fft3_lds(a.nf[axis]), s, P, in, out, conj);
  };
  Fft3Pass P{};
  if (to_grid) {
    // modes [n2][n1][n0] -> x: [n2][nf0][n1] -> y: [nf1][nf0][n2] -> z: grid [nf2][nf1][nf0]
    P = Fft3Pass{};
    P.nin = (int)n0; P.nout = (int)f0; P.in_modes = 1; P.out_modes = 0;
    P.nseq = n1 * n2; P.D1 = n1; P.P1 = n1 * n0; P.P2 = n0; P.Q1 = f0 * n1; P.Q2 = 1; P.ostride = n1;
    P.in_t = mode_t; P.out_t = mid_t; P.facA = fac[0]; P.facB = fac[1]; P.facC = fac[2];
    pass(0, P, static_cast<const float2*>(src), fp.mid);
    P = Fft3Pass{};
    P.nin = (int)n1; P.nout = (int)f1; P.in_modes = 1; P.out_modes = 0;
    P.nseq = f0 * n2; P.D1 = n2; P.P1 = n1; P.P2 = f0 * n1; P.Q1 = n2; P.Q2 = 1; P.ostride = f0 * n2;
    P.in_t = mid_t; P.out_t = mid2_t;
    pass(1, P, fp.mid, fp.mid2);
    P = Fft3Pass{};
    P.nin = (int)n2; P.nout = (int)f2; P.in_modes = 1; P.out_modes = 0;
    P.nseq = f0 * f1; P.D1 = f0; P.P1 = f0 * n2; P.P2 = n2; P.Q1 = f0; P.Q2 = 1; P.ostride = f0 * f1;
    P.in_t = mid2_t; P.out_t = grid_t;
    pass(2, P, fp.mid2, static_cast<float2*>(dst));
  } else {
    // grid [nf2][nf1][nf0] -> x: [nf2][n0][nf1] -> y: [n1][n0][nf2] -> z: modes [n2][n1][n0]
    P = Fft3Pass{};
    P.nin = (int)f0; P.nout = (int)n0; P.in_modes = 0; P.out_modes = 1;
    P.nseq = f1 * f2; P.D1 = f1; P.P1 = f1 * f0; P.P2 = f0; P.Q1 = n0 * f1; P.Q2 = 1; P.ostride = f1;
    P.in_t = grid_t; P.out_t = mid_t;
    pass(0, P, static_cast<const float2*>(src), fp.mid);
    P = Fft3Pass{};
    P.nin = (int)f1; P.nout = (int)n1; P.in_modes = 0; P.out_modes = 1;
    P.nseq = n0 * f2; P.D1 = f2; P.P1 = f1; P.P2 = n0 * f1; P.Q1 = f2; P.Q2 = 1; P.ostride = n0 * f2;
    P.in_t = mid_t; P.out_t = mid2_t;
    pass(1, P, fp.mid, fp.mid2);
    P = Fft3Pass{};
    P.nin = (int)f2; P.nout = (int)n2; P.in_modes = 0; P.out_modes = 1;
    P.nseq = n0 * n1; P.D1 = n0; P.P1 = n0 * f2; P.P2 = f2; P.Q1 = n0; P.Q2 = 1; P.ostride = n0 * n1;
    P.in_t = mid2_t; P.out_t = mode_t; P.facA = fac[2]; P.facB = fac[0]; P.facC = fac[1];
    pass(2, P, fp.mid2, static_cast<float2*>(dst));
  }
  return hipGetLastError();
}

// 1D: one pass per transform (one sequence; the k_fft3_pass block's other slots stay idle):
// grid -> kept modes times fac (type 1), or modes times fac at their bins -> grid (type 2)
hipError_t fft1_run(Fft2Plan& fp, const DeconvArgs& a, const void* src, void* dst, int isign, bool to_grid,
                    hipStream_t s) {
  const int64_t n = a.nm[0], nf = a.nf[0];
  const int logn = log2_exact(a.nf[0]);
  Fft3Pass P{};
  P.nin = (int)(to_grid ? n : nf);
  P.nout = (int)(to_grid ? nf : n);
  P.in_modes = to_grid ? 1 : 0;
  P.out_modes = to_grid ? 0 : 1;
  P.nmode = (int)n;
  P.modeord = a.modeord;
  P.nseq = 1;
  P.D1 = 1;
  P.ostride = 1;
  P.in_t = to_grid ? n : nf;
  P.out_t = to_grid ? nf : n;
  P.facA = static_cast<const float*>(a.fac[0]);
  P.tw = static_cast<const float2*>(fp.tw[0]);
  launch_logn<k_fft3_pass_l>(logn, dim3(1, (unsigned)a.n_trans), fft3_lds(a.nf[0]), s, P,
                             static_cast<const float2*>(src), static_cast<float2*>(dst), isign > 0 ? 1 : 0);
  return hipGetLastError();
}

hipError_t fft2_shuffle(Fft2Plan& fp, const DeconvArgs& a, const void* grid, void* f, int isign, hipStream_t s) {
  if (a.dim == 1) return fft1_run(fp, a, grid, f, isign, false, s);
  if (a.dim == 3) {
    hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * fft3_mid_cells(a), fp.retired);
    if (e == hipSuccess) e = grow_buffer(fp.mid2, fp.cap_mid2, a.n_trans * fft3_mid2_cells(a), fp.retired);
    if (e != hipSuccess) return e;
    return fft3_run(fp, a, grid, f, isign, false, s);
  }
  const Fft2K k = make_k(fp, a);
  hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * k.mid_cells, fp.retired);
  if (e != hipSuccess) return e;
  const int conj = isign > 0 ? 1 : 0;
  const unsigned nt = (unsigned)a.n_trans;
  launch_logn<k_fft2_rows_t1_l>(k.logn[0], dim3(fft_blocks(a.nf[1], a.nf[0]), nt), fft_lds(a.nf[0]), s, k,
                                static_cast<const float2*>(grid), fp.mid, conj);
  launch_logn<k_fft2_cols_t1_l>(k.logn[1], dim3(fft_blocks(a.nm[0], a.nf[1]), nt), fft_lds(a.nf[1]), s, k,
                                (const float2*)fp.mid, static_cast<float2*>(f), conj);
  return hipGetLastError();
}

hipError_t fft2_pad(Fft2Plan& fp, const DeconvArgs& a, const void* f, void* grid, int isign, hipStream_t s) {
  if (a.dim == 1) return fft1_run(fp, a, f, grid, isign, true, s);
  if (a.dim == 3) {
    hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * fft3_mid_cells(a), fp.retired);
    if (e == hipSuccess) e = grow_buffer(fp.mid2, fp.cap_mid2, a.n_trans * fft3_mid2_cells(a), fp.retired);
    if (e != hipSuccess) return e;
    return fft3_run(fp, a, f, grid, isign, true, s);
  }
  const Fft2K k = make_k(fp, a);
  hipError_t e = grow_buffer(fp.mid, fp.cap_mid, a.n_trans * k.mid_cells, fp.retired);
  if (e != hipSuccess) return e;
  const int conj = isign > 0 ? 1 : 0;
  const unsigned nt = (unsigned)a.n_trans;
  launch_logn<k_fft2_cols_t2_l>(k.logn[1], dim3(fft_blocks(a.nm[0], a.nf[1]), nt), fft_lds(a.nf[1]), s, k,
                                static_cast<const float2*>(f), fp.mid, conj);
  launch_logn<k_fft2_rows_t2_l>(k.logn[0], dim3(fft_blocks(a.nf[1], a.nf[0]), nt), fft_lds(a.nf[0]), s, k,
                                (const float2*)fp.mid, static_cast<float2*>(grid), conj);
  return hipGetLastError();
}

}  // namespace nufftx
