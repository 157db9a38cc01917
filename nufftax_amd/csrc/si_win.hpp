// 2D FP32 spreading on the matrix cores, one 16 x 16 window per sub-bin.
//
// The contribution of points j to the fine grid is a sum of outer products
//   fw[y][x] += sum_j (c_j wy_j(y)) wx_j(x)              (nufftax/core/spread.py:465-489)
// so the points of a small region accumulate as a matrix product A B with A the
// (rows x points) matrix of c_j wy_j and B the (points x columns) matrix of wx_j.
// A bin of 4 x 4 sub-bins of SB = 16 - ns support starts is one subproblem: the
// supports of the points of a sub-bin lie in a 16 x 16 window, so every group of
// 4 points of one sub-bin is one v_mfma_f32_16x16x4f32 per part (re, im) into a
// window accumulator held in registers; no per-point LDS atomics.
//
// Block = one subproblem (<= 2 * blockDim points of one bin):
//   1. stage: every point's support start gives its sub-bin; the points are counted
//      per sub-bin (LDS), and written into LDS in sub-bin order, each sub-bin padded to
//      a multiple of 4 with zero-strength entries, as {Zx, Zy, re c, im c} with
//      Z = round((i0 - xs - o) 2^23) the support offset relative to the sub-bin's window
//      (o = support start - window start, in [0, SB]) in fixed point;
//   2. accumulate: the groups of 4 are split into equal contiguous ranges, one per
//      wave.  Lane (i, k) of a group evaluates wy_k at window row i and wx_k at window
//      column i directly (z = (Z + i 2^23) 2^-23, ES kernel in registers, zero outside
//      the support) and feeds A[i][k] = c_k wy_k(i), B[k][i] = wx_k(i).  When the
//      wave's walk leaves a sub-bin, the window is added into the block's tile in
//      exact 64-bit fixed point (integer LDS atomics, 4 rows x 16 cells per step);
//   3. flush: the tile's non-zero cells go to the fine grid with FP32 global atomics.
// The fixed-point scale is 2^(50 - e) with 2^e above 2 n max(|re c|, |im c|) over the
// subproblem's n points (a block maximum taken while staging): every window value and
// partial tile sum stays below 2^50.  Windows accumulate in FP32 (like the reference's
// segment_sum); the tile sum of the windows is exact.
//
// MFMA operand layout (v_mfma_f32_16x16x4f32, 64 lanes):
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k,  D[i][j] in lane j + 16 (i / 4), element i % 4.
#pragma once

#include <cstdlib>

#include "si_kernels.hpp"

namespace nufftx {

typedef float win_f4 __attribute__((ext_vector_type(4)));

constexpr int kWinThreads = 512;  // 8 waves per subproblem
constexpr int kWinPpt = 2;        // points per thread while staging: subproblems of <= 1024 points
constexpr double kWinZScale = 8388608.0;  // 2^23

template <int NS>
struct WinGeom {
  static constexpr int SB = 16 - NS;  // sub-bin side (support starts); window 16 = SB + NS
  static constexpr int BS = 4 * SB;   // bin side
  static constexpr int T = BS + NS;   // tile side
};

// LDS bytes of the window spread for subproblems of maxsub points
template <int NS>
inline size_t win_spread_lds(int maxsub) {
  constexpr int T = WinGeom<NS>::T;
  return 2 * (size_t)T * T * sizeof(unsigned long long) + (size_t)(maxsub + 3 * 16) * sizeof(float4);
}

// ES kernel at z' = z 2^23 (c23 = c 2^-46): the same function as es_phi, zero outside the support
__device__ __forceinline__ float es_phi_scaled(float zu, float beta_log2e, float c23) {
  const float cz2 = c23 * zu * zu;
  float a = 1.0f - cz2;
  a = a > 0.0f ? a : 0.0f;
  const float s = __builtin_sqrtf(a);
  const float v = __builtin_amdgcn_exp2f(-beta_log2e * cz2 * __builtin_amdgcn_rcpf(1.0f + s));
  return cz2 <= 1.0f ? v : 0.0f;
}

template <typename CT, int NS, bool REC>
__global__ void __launch_bounds__(kWinThreads) k_spread_win2d(SIArgs<float> a, const cplx<float>* __restrict__ c,
                                                              cplx<float>* __restrict__ fw) {
  using G = WinGeom<NS>;
  using Word = unsigned long long;
  constexpr int SB = G::SB, T = G::T, TC = T * T;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Word* sre = reinterpret_cast<Word*>(smem_raw);
  Word* sim = sre + TC;
  int4* stage = reinterpret_cast<int4*>(sim + TC);
  __shared__ int cnt[16], fill[16], start[17];
  __shared__ unsigned smax;
  __shared__ int sbad;
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int4 sd = a.subprob[sub];
  const int n = sd.z;
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  for (int e = tid; e < TC; e += kWinThreads) {
    sre[e] = 0;
    sim[e] = 0;
  }
  if (tid < 16) cnt[tid] = 0;
  if (tid == 0) {
    smax = 0u;
    sbad = 0;
  }
  __syncthreads();

  // ---- 1. stage
  const cplx<float>* cT = c + (int64_t)tr * a.M;
  int4 rec[kWinPpt];
  int sb[kWinPpt];
  unsigned mymax = 0;
  bool bad = false;
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r) {
    const int j = tid + r * kWinThreads;
    sb[r] = -1;
    if (j >= n) continue;
    const int p = a.perm[sd.y + j];
    if ((unsigned)p >= (unsigned)a.M) continue;
    int l[2];
    double t[2];
    if constexpr (REC) {
      const uint2 w = a.crec[sd.y + j];
      l[0] = (int)(w.x >> 25);
      l[1] = (int)(w.y >> 25);
      t[0] = (double)(w.x & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1);
      t[1] = (double)(w.y & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1);
    } else {
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const Folded f = fold_coord((double)static_cast<const CT*>(a.coord[d])[p], a.nf[d]);
        const int i0 = support_start(f, NS / 2);
        l[d] = i0 - org[d];
        t[d] = (double)i0 - f.xs;
      }
    }
    int s2[2], Z[2];
#pragma unroll
    for (int d = 0; d < 2; ++d) {
      s2[d] = min(max(l[d], 0) / SB, 3);
      const double z = (t[d] - (double)(l[d] - s2[d] * SB)) * kWinZScale;
      bad = bad || !(fabs(z) < 1e9);  // non-finite coordinate
      Z[d] = bad ? 0 : (int)rint(z);
    }
    const cplx<float> cv = cT[p];
    mymax = max(mymax, max(__float_as_uint(fabsf(cv.re)), __float_as_uint(fabsf(cv.im))));
    rec[r] = make_int4(Z[0], Z[1], __float_as_int(cv.re), __float_as_int(cv.im));
    sb[r] = s2[1] * 4 + s2[0];
    atomicAdd(&cnt[sb[r]], 1);
  }
  // block maximum of |re c|, |im c| (non-negative floats order as their bit patterns;
  // a NaN pattern is above every finite one)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mymax = max(mymax, (unsigned)__shfl_xor((int)mymax, off, 64));
  if (lane == 0) atomicMax(&smax, mymax);
  if (bad) sbad = 1;
  __syncthreads();
  if (tid == 0) {  // padded sub-bin starts (multiples of 4)
    int acc = 0;
    for (int s = 0; s < 16; ++s) {
      start[s] = acc;
      fill[s] = acc;
      acc += (cnt[s] + 3) & ~3;
    }
    start[16] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r) {
    if (sb[r] < 0) continue;
    const int slot = atomicAdd(&fill[sb[r]], 1);
    stage[slot] = rec[r];
  }
  // zero-strength padding entries: z at the window origin, c = 0
  if (tid < 16 * 3) {
    const int s = tid / 3, q = tid % 3;
    const int slot = start[s] + cnt[s] + q;
    if (slot < start[s + 1]) stage[slot] = make_int4(0, 0, 0, 0);
  }
  const float cm = __uint_as_float(smax);
  const double B = 2.0 * (double)n * (double)cm;
  const bool finite = isfinite(B) && sbad == 0;
  double scale = 1.0, inv_scale = 1.0;
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);  // B < 2^e
    scale = ldexp(1.0, 50 - e);
    inv_scale = ldexp(1.0, e - 50);
  }
  __syncthreads();

  // ---- 2. accumulate
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr int W = kWinThreads / 64;
  const int NG = start[16] >> 2;
  const int g0 = (int)(((int64_t)NG * wave) / W), g1 = (int)(((int64_t)NG * (wave + 1)) / W);
  const int i = lane & 15, k = lane >> 4;
  const int ioff = i << 23;
  const float bl2e = a.es.beta_log2e;
  const float c23 = (float)((double)a.es.c / (kWinZScale * kWinZScale));
  win_f4 acc_re = {0.f, 0.f, 0.f, 0.f}, acc_im = {0.f, 0.f, 0.f, 0.f};
  int s = 0;
  while (s < 15 && start[s + 1] <= 4 * g0) ++s;
  auto flush = [&](int sbin) {
    const int x = (sbin & 3) * SB + i;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int cell = ((sbin >> 2) * SB + 4 * k + rr) * T + x;
      if (acc_re[rr] != 0.0f) atomicAdd(&sre[cell], TileAcc<float>::quant((double)acc_re[rr], scale));
      if (acc_im[rr] != 0.0f) atomicAdd(&sim[cell], TileAcc<float>::quant((double)acc_im[rr], scale));
    }
  };
  for (int g = g0; g < g1 && finite; ++g) {
    if (4 * g >= start[s + 1]) {
      flush(s);
      acc_re = win_f4{0.f, 0.f, 0.f, 0.f};
      acc_im = win_f4{0.f, 0.f, 0.f, 0.f};
      do {
        ++s;
      } while (s < 15 && start[s + 1] <= 4 * g);
    }
    const int4 rc = stage[4 * g + k];
    const float wy = es_phi_scaled((float)(rc.y + ioff), bl2e, c23);
    const float wx = es_phi_scaled((float)(rc.x + ioff), bl2e, c23);
    const float are = __int_as_float(rc.z) * wy, aim = __int_as_float(rc.w) * wy;
    acc_re = __builtin_amdgcn_mfma_f32_16x16x4f32(are, wx, acc_re, 0, 0, 0);
    acc_im = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, wx, acc_im, 0, 0, 0);
  }
  if (g1 > g0 && finite) flush(s);
  __syncthreads();

  // ---- 3. flush the tile's non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < TC; cell += kWinThreads) {
    const long long ir = (long long)sre[cell], ii = (long long)sim[cell];
    if (finite && ir == 0 && ii == 0) continue;
    const float vr = finite ? (float)((double)ir * inv_scale) : NAN;
    const float vi = finite ? (float)((double)ii * inv_scale) : NAN;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// NUFFTX_SPREAD_WIN=1 selects the window spread where it applies
inline bool use_win_spread() {
  const char* e = std::getenv("NUFFTX_SPREAD_WIN");
  return e != nullptr && e[0] == '1';
}

// Whether the window spread applies to a geometry: 2D FP32, ns <= 8 (window 16 = SB + ns
// with SB >= 8), square bins of 4 x 4 sub-bins, subproblems of at most 2 points per thread.
inline bool win2d_applicable(const Geometry& g, int prec, int maxsub) {
  if (g.dim != 2 || prec != kF32 || g.ns > 8 || maxsub > kWinPpt * kWinThreads) return false;
  const int bs = 4 * (16 - g.ns);
  return g.bs[0] == bs && g.bs[1] == bs && g.tile[0] == bs + g.ns && g.tile[1] == bs + g.ns;
}

}  // namespace nufftx
