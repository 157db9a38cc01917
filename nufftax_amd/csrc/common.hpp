// Shared device helpers for the nufftx HIP library (gfx950).
//
// Conventions follow nufftax (GragasLab/nufftax):
//   * fold: xs = ((x / 2pi + 1/2) mod 1) * nf in [0, nf)   (nufftax/core/spread.py:78-97)
//   * support start i0 = ceil(xs - ns/2), z_k = i0 + k - xs  (spread.py:216-230)
//   * ES kernel phi(z) = exp(beta (sqrt(1 - c z^2) - 1))    (nufftax/core/kernel.py:79-111)
//   * grids are (n_trans, nf3, nf2, nf1), x fastest          (spread.py:466, 632-634)
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nufftx {

constexpr int kMaxDim = 3;
// 1 / (2 pi) formed exactly as the reference forms it: 1.0 / (2.0 * pi) in double.
constexpr double kInv2Pi = 1.0 / (2.0 * 3.141592653589793);
constexpr double kTwoPi = 2.0 * 3.141592653589793;

template <typename T>
struct cplx {
  T re, im;
};

// ---------------------------------------------------------------------------
// FP64 fold of one coordinate.  Multiply and add are separately rounded
// (__dmul_rn / __dadd_rn) so that the bin of every point is a pure function of
// its coordinate, identical in every kernel and in a NumPy restatement.
// Non-finite coordinates fold to 0 for binning; their weights become NaN.
// ---------------------------------------------------------------------------
struct Folded {
  double xs;  // in [0, nf), or NaN for a non-finite input
  int cell;   // floor(xs), 0 for non-finite input
};

__device__ __forceinline__ Folded fold_coord(double x, int nf) {
  double r = __dadd_rn(__dmul_rn(x, kInv2Pi), 0.5);
  r = r - floor(r);
  double xs = r * (double)nf;
  if (xs >= (double)nf) xs -= (double)nf;
  Folded f;
  f.xs = xs;
  if (xs >= 0.0 && xs < (double)nf) {
    f.cell = (int)xs;
  } else {
    f.cell = 0;
  }
  return f;
}

// i0 = ceil(xs - ns/2) as in spread.py:216, computed in FP64; for NaN input a
// valid in-bin value derived from the (sanitised) cell is returned instead.
__device__ __forceinline__ int support_start(const Folded& f, int half) {
  if (f.xs >= 0.0) return (int)ceil(f.xs - (double)half);
  return f.cell - half;
}

// ---------------------------------------------------------------------------
// ES kernel.  exp(beta (s - 1)) with s = sqrt(1 - c z^2) is evaluated as
// exp(-beta c z^2 / (1 + s)) to avoid the cancellation in s - 1; both forms
// are the same function.  The float path uses the hardware sqrt, reciprocal and
// exp2 (v_sqrt_f32 / v_rcp_f32 / v_exp_f32, about 1 ulp each): the correctly
// rounded sqrtf is a 17-instruction sequence on gfx950 (4.9 lane-ops per clock
// per CU against 22 for v_exp_f32), and the kernel's value does not need it --
// a 1-ulp error in s moves phi by under 1e-8 of its peak.
// ---------------------------------------------------------------------------
struct EsConst {
  double beta;
  double c;
};

__device__ __forceinline__ float es_phi(float z, float beta_log2e, float c) {
  float cz2 = c * z * z;
  float a = 1.0f - cz2;
  a = a > 0.0f ? a : 0.0f;
  float s = __builtin_amdgcn_sqrtf(a);
  float e = -beta_log2e * cz2 * __builtin_amdgcn_rcpf(1.0f + s);
  float v = __builtin_amdgcn_exp2f(e);
  return cz2 <= 1.0f ? v : 0.0f;
}

// es_phi for arguments inside the support (c z^2 <= 1 up to rounding): the same value there,
// without the range select (outside, the clamp gives exp(-beta c z^2) instead of 0)
__device__ __forceinline__ float es_phi_in(float z, float beta_log2e, float c) {
  const float cz2 = c * z * z;
  const float a = fmaxf(1.0f - cz2, 0.0f);
  const float s = __builtin_amdgcn_sqrtf(a);
  return __builtin_amdgcn_exp2f(-beta_log2e * cz2 * __builtin_amdgcn_rcpf(1.0f + s));
}

// ---------------------------------------------------------------------------
// Polynomial ES kernel for ns = 8 (FP32 kernels).  The weight of support interval k of a point
// (z = t + k, t = i0 - xs in [-ns/2, 1 - ns/2]) is phi(z) = P_k(v), v = 2 (t + ns/2) - 1 in
// [-1, 1], one v for all k: the piecewise-polynomial kernel evaluation of FINUFFT (Barnett,
// Magland, af Klinteberg 2019).  phi is even, so P_{ns-1-k}(v) = P_k(-v), and with
// P_k(v) = E_k(v^2) + v O_k(v^2) (degree 9: E_k and O_k of degree 4) a pair of weights costs
// 10 FMAs: w[k] = E + v O, w[ns-1-k] = E - v O; 40 FP32 coefficients in all, against 7 plain
// and 3 transcendental instructions (sqrt, rcp, exp2) per weight for es_phi_in.  The host
// interpolates P_k at Chebyshev points in FP64 (es_poly_table), evaluates the FP32 form as the
// kernels do at 257 points per interval against the exact phi, and switches it on only where
// the largest error is at most kEsPolyTol (phi(0) = 1): at upsampfac 2, 8e-8 (the FP32
// rounding of values near 1) on the inner intervals and 5e-9 on the two outer ones, where the
// square root's branch point at |z| = ns/2 stops the convergence near exp(-beta).
// ---------------------------------------------------------------------------
constexpr int kEsPolyNs = 8;
constexpr double kEsPolyTol = 2.5e-7;
struct EsPoly {
  float e[kEsPolyNs / 2][5];  // E_k(u) = sum_d e[k][d] u^d
  float o[kEsPolyNs / 2][5];  // O_k(u) = sum_d o[k][d] u^d
  int on;                     // 1: the FP32 kernels take their ns = 8 weights from the polynomials
};

// w[k] = phi(t + k) for k < NS = 8, from v = 2 (t + NS/2) - 1 (coefficients uniform over the wave)
template <int NS>
__device__ __forceinline__ void es_poly_weights(float v, const EsPoly& p, float (&w)[NS]) {
  static_assert(NS == kEsPolyNs, "the polynomial ES kernel is fitted for ns = 8");
  const float u = v * v;
#pragma unroll
  for (int k = 0; k < NS / 2; ++k) {
    float ev = p.e[k][4], od = p.o[k][4];
#pragma unroll
    for (int d = 3; d >= 0; --d) {
      ev = fmaf(ev, u, p.e[k][d]);
      od = fmaf(od, u, p.o[k][d]);
    }
    w[k] = fmaf(v, od, ev);
    w[NS - 1 - k] = fmaf(-v, od, ev);
  }
}

// v of a compact record word (l << 25 | round((t + ns/2 + 1) 2^24)): 2 (t + ns/2) - 1
__device__ __forceinline__ float es_poly_v(unsigned word) {
  return (float)((double)(word & 0x1FFFFFFu) * (1.0 / 8388608.0) - 3.0);
}

__device__ __forceinline__ double es_phi(double z, double beta, double c) {
  double cz2 = c * z * z;
  double a = 1.0 - cz2;
  a = a > 0.0 ? a : 0.0;
  double s = sqrt(a);
  double v = exp(-beta * cz2 / (1.0 + s));
  return cz2 <= 1.0 ? v : 0.0;
}

// phi and dphi/dz (kernel.py:114-162: derivative zero where 1 - c z^2 <= 1e-14).
__device__ __forceinline__ void es_phi_dphi(float z, float beta_log2e, float beta, float c, float& phi,
                                            float& dphi) {
  float cz2 = c * z * z;
  float a = 1.0f - cz2;
  float s = __builtin_amdgcn_sqrtf(a > 1e-14f ? a : 1e-14f);
  float v = __builtin_amdgcn_exp2f(-beta_log2e * cz2 * __builtin_amdgcn_rcpf(1.0f + s));
  phi = a >= 0.0f ? v : 0.0f;
  dphi = a > 1e-14f ? -beta * c * z * __builtin_amdgcn_rcpf(s) * v : 0.0f;
}

__device__ __forceinline__ void es_phi_dphi(double z, double /*beta_log2e*/, double beta, double c, double& phi,
                                            double& dphi) {
  double cz2 = c * z * z;
  double a = 1.0 - cz2;
  double s = sqrt(a > 1e-14 ? a : 1e-14);
  double v = exp(-beta * cz2 / (1.0 + s));
  phi = a >= 0.0 ? v : 0.0;
  dphi = a > 1e-14 ? -beta * c * z / s * v : 0.0;
}

template <typename T>
struct EsArgs {
  T beta_log2e;  // beta * log2(e) (float path) or beta (double path)
  T beta;
  T c;
};

// Weights of one dimension for a point: w[k] = phi(t + k), t = i0 - xs.
template <typename T, int NS>
__device__ __forceinline__ void es_weights(T t, const EsArgs<T>& ea, T (&w)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) w[k] = es_phi(t + (T)k, ea.beta_log2e, ea.c);
}

template <typename T, int NS>
__device__ __forceinline__ void es_weights_deriv(T t, const EsArgs<T>& ea, T (&w)[NS], T (&dw)[NS]) {
#pragma unroll
  for (int k = 0; k < NS; ++k) es_phi_dphi(t + (T)k, ea.beta_log2e, ea.beta, ea.c, w[k], dw[k]);
}

// ---------------------------------------------------------------------------
// Wave64 helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

// number of set bits of `mask` in lanes below the calling lane
__device__ __forceinline__ int popc_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0));
}

template <typename T>
__device__ __forceinline__ void atomic_add_global(T* p, T v) {
  unsafeAtomicAdd(p, v);
}

}  // namespace nufftx
