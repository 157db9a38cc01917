// Elementwise stages of the pipeline:
//   * fill kernel (used instead of hipMemsetAsync so captured graphs replay correctly)
//   * fused deconvolve + mode truncation + shift for type 1
//     (nufftax/core/deconvolve.py:81-111, 174-224, 312-366)
//   * fused deconvolve + zero padding for type 2 (deconvolve.py:114-166, 227-304, 369-468)
//   * type-3 FP64 centring/rescale/phase kernels (nufftax/transforms/nufft3.py:257-299,
//     347-394, 444-510)
#include <cmath>

#include "internal.hpp"

namespace nufftx {

namespace {

__global__ void k_zero(uint4* __restrict__ p, int64_t n16, unsigned char* __restrict__ tail, int ntail) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) p[i] = make_uint4(0, 0, 0, 0);
  if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

unsigned grid_for(int64_t n, int threads = 256, int64_t cap = 1 << 16) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// mode index m (output order) -> fine-grid index (deconvolve.py:59-73)
__device__ __forceinline__ int mode_to_grid(int m, int n, int nf, int modeord) {
  const int kmin = -(n / 2), kmax = (n - 1) / 2;
  int k;
  if (modeord == 0)
    k = m + kmin;
  else
    k = m <= kmax ? m : m - n;
  return k >= 0 ? k : nf + k;
}

// fine-grid index -> mode index, or -1 if the cell is not a kept mode
__device__ __forceinline__ int grid_to_mode(int g, int n, int nf, int modeord) {
  const int kmin = -(n / 2), kmax = (n - 1) / 2;
  int k;
  if (g <= kmax)
    k = g;
  else if (g >= nf + kmin)
    k = g - nf;
  else
    return -1;
  if (modeord == 0) return k - kmin;
  return k >= 0 ? k : n + k;
}

template <typename T>
struct DK {
  int64_t n_trans;
  int nf[3];
  int nm[3];
  int modeord;
  const T* fac[3];  // factors in output mode order, or (PH) phihat[0 .. n/2]
};

// Deconvolution factor of mode index m on axis d: a table lookup, or (PH) the
// factor (-1)^k / phihat[|k|] of that mode (deconvolve.py:27-56) derived in place.
template <bool PH, typename T>
__device__ __forceinline__ T dfac(const DK<T>& a, int d, int m) {
  if (!PH) return a.fac[d][m];
  const int n = a.nm[d];
  const int k = a.modeord == 0 ? m - n / 2 : (m <= (n - 1) / 2 ? m : m - n);
  const int ak = k < 0 ? -k : k;
  const T inv = (T)1 / a.fac[d][ak];
  return (ak & 1) ? -inv : inv;
}

template <typename T, int D, bool PH>
__global__ void __launch_bounds__(256) k_deconv_shuffle(DK<T> a, const cplx<T>* __restrict__ in,
                                                        cplx<T>* __restrict__ out) {
  const int64_t nm = (int64_t)a.nm[0] * a.nm[1] * a.nm[2];
  const int64_t ng = (int64_t)a.nf[0] * a.nf[1] * a.nf[2];
  const int64_t total = nm * a.n_trans;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const int m0 = (int)(idx % a.nm[0]);
    int64_t r = idx / a.nm[0];
    const int m1 = D >= 2 ? (int)(r % a.nm[1]) : 0;
    if (D >= 2) r /= a.nm[1];
    const int m2 = D >= 3 ? (int)(r % a.nm[2]) : 0;
    if (D >= 3) r /= a.nm[2];
    const int64_t tr = r;
    int64_t gidx = mode_to_grid(m0, a.nm[0], a.nf[0], a.modeord);
    T fac = dfac<PH>(a, 0, m0);
    if (D >= 2) {
      gidx += (int64_t)mode_to_grid(m1, a.nm[1], a.nf[1], a.modeord) * a.nf[0];
      fac *= dfac<PH>(a, 1, m1);
    }
    if (D >= 3) {
      gidx += (int64_t)mode_to_grid(m2, a.nm[2], a.nf[2], a.modeord) * a.nf[0] * a.nf[1];
      fac *= dfac<PH>(a, 2, m2);
    }
    const cplx<T> v = in[tr * ng + gidx];
    out[idx] = {v.re * fac, v.im * fac};
  }
}

template <typename T, int D, bool PH>
__global__ void __launch_bounds__(256) k_deconv_pad(DK<T> a, const cplx<T>* __restrict__ in,
                                                    cplx<T>* __restrict__ out) {
  const int64_t nm = (int64_t)a.nm[0] * a.nm[1] * a.nm[2];
  const int64_t ng = (int64_t)a.nf[0] * a.nf[1] * a.nf[2];
  const int64_t total = ng * a.n_trans;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const int g0 = (int)(idx % a.nf[0]);
    int64_t r = idx / a.nf[0];
    const int g1 = D >= 2 ? (int)(r % a.nf[1]) : 0;
    if (D >= 2) r /= a.nf[1];
    const int g2 = D >= 3 ? (int)(r % a.nf[2]) : 0;
    if (D >= 3) r /= a.nf[2];
    const int64_t tr = r;
    cplx<T> v = {0, 0};
    const int m0 = grid_to_mode(g0, a.nm[0], a.nf[0], a.modeord);
    const int m1 = D >= 2 ? grid_to_mode(g1, a.nm[1], a.nf[1], a.modeord) : 0;
    const int m2 = D >= 3 ? grid_to_mode(g2, a.nm[2], a.nf[2], a.modeord) : 0;
    if (m0 >= 0 && m1 >= 0 && m2 >= 0) {
      T fac = dfac<PH>(a, 0, m0);
      int64_t midx = m0;
      if (D >= 2) {
        fac *= dfac<PH>(a, 1, m1);
        midx += (int64_t)m1 * a.nm[0];
      }
      if (D >= 3) {
        fac *= dfac<PH>(a, 2, m2);
        midx += (int64_t)m2 * a.nm[0] * a.nm[1];
      }
      const cplx<T> f = in[tr * nm + midx];
      v = {f.re * fac, f.im * fac};
    }
    out[idx] = v;
  }
}

template <typename T>
DK<T> make_dk(const DeconvArgs& a) {
  DK<T> k;
  k.n_trans = a.n_trans;
  k.modeord = a.modeord;
  for (int d = 0; d < 3; ++d) {
    k.nf[d] = d < a.dim ? a.nf[d] : 1;
    k.nm[d] = d < a.dim ? a.nm[d] : 1;
    k.fac[d] = static_cast<const T*>(a.fac[d]);
  }
  return k;
}

template <typename T>
hipError_t shuffle_t(const DeconvArgs& a, const void* in, void* out, hipStream_t s) {
  const DK<T> k = make_dk<T>(a);
  const int64_t total = (int64_t)k.nm[0] * k.nm[1] * k.nm[2] * k.n_trans;
  if (total == 0) return hipSuccess;
  const unsigned g = grid_for(total);
  auto i = static_cast<const cplx<T>*>(in);
  auto o = static_cast<cplx<T>*>(out);
  const bool ph = a.from_phihat;
  if (a.dim == 1) {
    if (ph) hipLaunchKernelGGL((k_deconv_shuffle<T, 1, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_shuffle<T, 1, false>), dim3(g), dim3(256), 0, s, k, i, o);
  } else if (a.dim == 2) {
    if (ph) hipLaunchKernelGGL((k_deconv_shuffle<T, 2, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_shuffle<T, 2, false>), dim3(g), dim3(256), 0, s, k, i, o);
  } else {
    if (ph) hipLaunchKernelGGL((k_deconv_shuffle<T, 3, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_shuffle<T, 3, false>), dim3(g), dim3(256), 0, s, k, i, o);
  }
  return hipGetLastError();
}

template <typename T>
hipError_t pad_t(const DeconvArgs& a, const void* in, void* out, hipStream_t s) {
  const DK<T> k = make_dk<T>(a);
  const int64_t total = (int64_t)k.nf[0] * k.nf[1] * k.nf[2] * k.n_trans;
  if (total == 0) return hipSuccess;
  const unsigned g = grid_for(total);
  auto i = static_cast<const cplx<T>*>(in);
  auto o = static_cast<cplx<T>*>(out);
  const bool ph = a.from_phihat;
  if (a.dim == 1) {
    if (ph) hipLaunchKernelGGL((k_deconv_pad<T, 1, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_pad<T, 1, false>), dim3(g), dim3(256), 0, s, k, i, o);
  } else if (a.dim == 2) {
    if (ph) hipLaunchKernelGGL((k_deconv_pad<T, 2, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_pad<T, 2, false>), dim3(g), dim3(256), 0, s, k, i, o);
  } else {
    if (ph) hipLaunchKernelGGL((k_deconv_pad<T, 3, true>), dim3(g), dim3(256), 0, s, k, i, o);
    else hipLaunchKernelGGL((k_deconv_pad<T, 3, false>), dim3(g), dim3(256), 0, s, k, i, o);
  }
  return hipGetLastError();
}

// ------------------------------------------------------------------ type 3
constexpr int kRedBlocks = 256;

__device__ __forceinline__ void wave_minmax(double& mn, double& mx) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, off, 64));
    mx = fmax(mx, __shfl_xor(mx, off, 64));
  }
}

template <typename CT>
__global__ void __launch_bounds__(256) k_minmax(const CT* __restrict__ p, int64_t n, double* __restrict__ out) {
  __shared__ double smn[4], smx[4];
  double mn = INFINITY, mx = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double v = (double)p[i];
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  wave_minmax(mn, mx);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    smn[w] = mn;
    smx[w] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < 4; ++k) {
      mn = fmin(mn, smn[k]);
      mx = fmax(mx, smx[k]);
    }
    out[2 * blockIdx.x] = mn;
    out[2 * blockIdx.x + 1] = mx;
  }
}

struct T3Fin {
  int dim;
  int nf[3];
  double upsampfac;
};

__global__ void k_t3_finalize(T3Fin f, const double* __restrict__ part, T3State* __restrict__ st) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int d = 0; d < 3; ++d) {
    if (d >= f.dim) {
      st->C[d] = st->D[d] = 0.0;
      st->gamma[d] = st->h[d] = 1.0;
      continue;
    }
    double xmn = INFINITY, xmx = -INFINITY, smn = INFINITY, smx = -INFINITY;
    for (int b = 0; b < kRedBlocks; ++b) {
      xmn = fmin(xmn, part[(2 * d) * kRedBlocks + 2 * b]);
      xmx = fmax(xmx, part[(2 * d) * kRedBlocks + 2 * b + 1]);
      smn = fmin(smn, part[(2 * (3 + d)) * kRedBlocks + 2 * b]);
      smx = fmax(smx, part[(2 * (3 + d)) * kRedBlocks + 2 * b + 1]);
    }
    if (!(xmn <= xmx)) xmn = xmx = 0.0;  // no sources
    if (!(smn <= smx)) smn = smx = 0.0;  // no targets
    const double C = (xmn + xmx) / 2.0;
    const double Dd = (smn + smx) / 2.0;
    const double S = fmax((smx - smn) / 2.0, 1e-10);
    st->C[d] = C;
    st->D[d] = Dd;
    st->h[d] = kTwoPi / (double)f.nf[d];
    st->gamma[d] = (double)f.nf[d] / (2.0 * f.upsampfac * S);
  }
}

template <typename CT>
__global__ void __launch_bounds__(256) k_t3_prep_src(int dim, int64_t M, const CT* __restrict__ x0,
                                                     const CT* __restrict__ x1, const CT* __restrict__ x2,
                                                     const T3State* __restrict__ st, int isign,
                                                     double* __restrict__ r0, double* __restrict__ r1,
                                                     double* __restrict__ r2, double* __restrict__ pre) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) {
    const CT* xs[3] = {x0, x1, x2};
    double* rs[3] = {r0, r1, r2};
    double ph = 0.0;
    for (int d = 0; d < dim; ++d) {
      const double x = (double)xs[d][i];
      rs[d][i] = (x - st->C[d]) / st->gamma[d];
      ph += st->D[d] * x;
    }
    double sn, cs;
    sincos((double)isign * ph, &sn, &cs);
    pre[2 * i] = cs;
    pre[2 * i + 1] = sn;
  }
}

struct QuadDev {
  int n;
  double z[40];
  double wphi[40];
};

template <typename CT>
__global__ void __launch_bounds__(256) k_t3_prep_tgt(int dim, int64_t N, const CT* __restrict__ s0,
                                                     const CT* __restrict__ s1, const CT* __restrict__ s2,
                                                     const T3State* __restrict__ st, int isign, QuadDev q,
                                                     double* __restrict__ r0, double* __restrict__ r1,
                                                     double* __restrict__ r2, double* __restrict__ post) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < N; k += stride) {
    const CT* ss[3] = {s0, s1, s2};
    double* rs[3] = {r0, r1, r2};
    double ph = 0.0, phi = 1.0;
    for (int d = 0; d < dim; ++d) {
      const double s = (double)ss[d][k];
      const double sr = st->h[d] * st->gamma[d] * (s - st->D[d]);
      rs[d][k] = sr;
      ph += (s - st->D[d]) * st->C[d];
      double acc = 0.0;
      for (int j = 0; j < q.n; ++j) acc += q.wphi[j] * cos(sr * q.z[j]);
      phi *= acc;
    }
    double sn, cs;
    sincos((double)isign * ph, &sn, &cs);
    post[2 * k] = cs / phi;
    post[2 * k + 1] = sn / phi;
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_cmul_pointwise(int64_t n_trans, int64_t M, const cplx<T>* __restrict__ in,
                                                        const double* __restrict__ fac, cplx<T>* __restrict__ out) {
  const int64_t total = n_trans * M;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
    const int64_t i = idx % M;
    const cplx<T> v = in[idx];
    const double fr = fac[2 * i], fi = fac[2 * i + 1];
    const double vr = (double)v.re, vi = (double)v.im;
    out[idx] = {(T)(vr * fr - vi * fi), (T)(vr * fi + vi * fr)};
  }
}

}  // namespace

hipError_t launch_zero(void* p, int64_t bytes, hipStream_t s) {
  if (bytes <= 0) return hipSuccess;
  const int64_t n16 = bytes / 16;
  const int ntail = (int)(bytes - n16 * 16);
  hipLaunchKernelGGL(k_zero, dim3(grid_for(std::max<int64_t>(n16, 1))), dim3(256), 0, s, static_cast<uint4*>(p), n16,
                     static_cast<unsigned char*>(p) + n16 * 16, ntail);
  return hipGetLastError();
}

hipError_t launch_deconv_shuffle(const DeconvArgs& a, const void* fw_hat, void* f, hipStream_t s) {
  return a.prec == kF32 ? shuffle_t<float>(a, fw_hat, f, s) : shuffle_t<double>(a, fw_hat, f, s);
}

hipError_t launch_deconv_pad(const DeconvArgs& a, const void* f, void* fw_hat, hipStream_t s) {
  return a.prec == kF32 ? pad_t<float>(a, f, fw_hat, s) : pad_t<double>(a, f, fw_hat, s);
}

hipError_t launch_t3_setup(int dim, int prec, int64_t M, const void* const x[3], int64_t N, const void* const sv[3],
                           const int nf[3], double upsampfac, T3State* st, double* scratch, int scratch_elems,
                           hipStream_t strm) {
  if (scratch_elems < 12 * kRedBlocks) return hipErrorInvalidValue;
  for (int d = 0; d < 3; ++d) {
    double* px = scratch + (2 * d) * kRedBlocks;
    double* ps = scratch + (2 * (3 + d)) * kRedBlocks;
    const int64_t mx = d < dim ? M : 0;
    const int64_t ns = d < dim ? N : 0;
    if (prec == kF32) {
      hipLaunchKernelGGL(k_minmax<float>, dim3(kRedBlocks), dim3(256), 0, strm, static_cast<const float*>(x[d]), mx,
                         px);
      hipLaunchKernelGGL(k_minmax<float>, dim3(kRedBlocks), dim3(256), 0, strm, static_cast<const float*>(sv[d]),
                         ns, ps);
    } else {
      hipLaunchKernelGGL(k_minmax<double>, dim3(kRedBlocks), dim3(256), 0, strm, static_cast<const double*>(x[d]),
                         mx, px);
      hipLaunchKernelGGL(k_minmax<double>, dim3(kRedBlocks), dim3(256), 0, strm, static_cast<const double*>(sv[d]),
                         ns, ps);
    }
  }
  T3Fin f;
  f.dim = dim;
  for (int d = 0; d < 3; ++d) f.nf[d] = nf[d];
  f.upsampfac = upsampfac;
  hipLaunchKernelGGL(k_t3_finalize, dim3(1), dim3(64), 0, strm, f, scratch, st);
  return hipGetLastError();
}

hipError_t launch_t3_prep_src(int dim, int prec, int64_t M, const void* const x[3], const T3State* st, int isign,
                              double* xr[3], double* pre, hipStream_t s) {
  if (M == 0) return hipSuccess;
  const unsigned g = grid_for(M);
  if (prec == kF32)
    hipLaunchKernelGGL(k_t3_prep_src<float>, dim3(g), dim3(256), 0, s, dim, M, static_cast<const float*>(x[0]),
                       static_cast<const float*>(x[1]), static_cast<const float*>(x[2]), st, isign, xr[0], xr[1],
                       xr[2], pre);
  else
    hipLaunchKernelGGL(k_t3_prep_src<double>, dim3(g), dim3(256), 0, s, dim, M, static_cast<const double*>(x[0]),
                       static_cast<const double*>(x[1]), static_cast<const double*>(x[2]), st, isign, xr[0], xr[1],
                       xr[2], pre);
  return hipGetLastError();
}

hipError_t launch_t3_prep_tgt(int dim, int prec, int64_t N, const void* const sv[3], const T3State* st, int isign,
                              const QuadRule& q, double* sr[3], double* post, hipStream_t strm) {
  if (N == 0) return hipSuccess;
  QuadDev qd;
  qd.n = q.n;
  for (int j = 0; j < 40; ++j) {
    qd.z[j] = q.z[j];
    qd.wphi[j] = q.wphi[j];
  }
  const unsigned g = grid_for(N);
  if (prec == kF32)
    hipLaunchKernelGGL(k_t3_prep_tgt<float>, dim3(g), dim3(256), 0, strm, dim, N, static_cast<const float*>(sv[0]),
                       static_cast<const float*>(sv[1]), static_cast<const float*>(sv[2]), st, isign, qd, sr[0],
                       sr[1], sr[2], post);
  else
    hipLaunchKernelGGL(k_t3_prep_tgt<double>, dim3(g), dim3(256), 0, strm, dim, N, static_cast<const double*>(sv[0]),
                       static_cast<const double*>(sv[1]), static_cast<const double*>(sv[2]), st, isign, qd, sr[0],
                       sr[1], sr[2], post);
  return hipGetLastError();
}

hipError_t launch_cmul_pointwise(int prec, int64_t n_trans, int64_t M, const void* in, const double* fac, void* out,
                                 hipStream_t s) {
  if (M == 0 || n_trans == 0) return hipSuccess;
  const unsigned g = grid_for(M * n_trans);
  if (prec == kF32)
    hipLaunchKernelGGL(k_cmul_pointwise<float>, dim3(g), dim3(256), 0, s, n_trans, M,
                       static_cast<const cplx<float>*>(in), fac, static_cast<cplx<float>*>(out));
  else
    hipLaunchKernelGGL(k_cmul_pointwise<double>, dim3(g), dim3(256), 0, s, n_trans, M,
                       static_cast<const cplx<double>*>(in), fac, static_cast<cplx<double>*>(out));
  return hipGetLastError();
}

}  // namespace nufftx
