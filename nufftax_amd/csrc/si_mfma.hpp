// 2D FP32 spreading on the matrix cores (gfx950 MFMA).
//
// The contribution of points j to a tile is a sum of outer products,
//   tile[y][x] += sum_j (c_j wy_j[y - y0_j]) * wx_j[x - x0_j]      (nufftax/core/spread.py:465-489)
// i.e. a product A B of a (rows x points) and a (points x cols) matrix whose
// columns/rows carry the ns nonzero kernel weights of each point.  A bin of
// 4 x 4 sub-bins of SB = 16 - ns cells is spread with 16x16 windows: every point
// of a sub-bin has its ns x ns support inside the sub-bin's 16 x 16 window, so one
// v_mfma_f32_16x16x4f32 accumulates 4 points into the window (4 x 16 A operand =
// c * y-weights, 16 x 4 B operand = x-weights).  Accumulators of all 16 sub-bins
// live in registers for the whole subproblem; LDS only stages per-point weights
// and the final tile.  Nothing is accumulated with LDS atomics.
//
// Work split: a block of kWaves waves takes one subproblem (<= maxsub points of
// one bin); wave w processes 64-point chunks w, w + kWaves, ...  Per chunk every
// lane folds one point (FP64), evaluates its 2 ns weights once, and the wave
// orders the chunk by sub-bin with 16 ballots.  Each wave then adds its
// accumulators into its own LDS tile; the block sums the wave tiles and flushes
// non-zero cells to the fine grid with global atomics.
//
// MFMA operand layout (v_mfma_f32_16x16x4f32, 64 lanes):
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k,
//   D[i][j] in lane j + 16 (i / 4), element i % 4.
#pragma once

#include "si_kernels.hpp"

namespace nufftx {

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaWaves = 2;

template <int NS>
struct MfmaGeom {
  static constexpr int SB = 16 - NS;       // sub-bin side (cells)
  static constexpr int NSUB = 4;           // sub-bins per bin side
  static constexpr int BS = NSUB * SB;     // bin side
  static constexpr int T = BS + NS;        // tile side
  static constexpr int TC = T * T;
  static constexpr int REC = 3 * NS + 1;   // staged floats per point: wx[NS], (c wy)[NS] complex, offsets
  static constexpr int WAVE_FLOATS = 2 * TC + 64 * REC;
  static constexpr size_t lds_bytes() { return (size_t)kMfmaWaves * WAVE_FLOATS * sizeof(float); }
};

template <typename CT, int NS>
__global__ void __launch_bounds__(64 * kMfmaWaves) k_spread_mfma2d(SIArgs<float> a, const cplx<float>* __restrict__ c,
                                                                    cplx<float>* __restrict__ fw) {
  using G = MfmaGeom<NS>;
  constexpr int SB = G::SB, T = G::T, TC = G::TC, REC = G::REC;
  extern __shared__ __align__(16) float smem_f[];
  const int wave = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  float* tre = smem_f + wave * G::WAVE_FLOATS;
  float* tim = tre + TC;
  float* stage = tim + TC;

  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const int bin_x0 = org[0] + NS / 2, bin_y0 = org[1] + NS / 2;  // first cell of the bin
  for (int i = lane; i < TC; i += 64) {
    tre[i] = 0.0f;
    tim[i] = 0.0f;
  }
  const cplx<float>* cT = c + (int64_t)tr * a.M;

  mfma_f4 acc_re[16], acc_im[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    acc_re[s] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    acc_im[s] = mfma_f4{0.f, 0.f, 0.f, 0.f};
  }
  const int kk = lane >> 4;  // point of the 4-point group this lane feeds
  const int rr = lane & 15;  // window row (A) / column (B) this lane feeds

  const int nchunk = (sd.z + 63) / 64;
  for (int ch = wave; ch < nchunk; ch += kMfmaWaves) {
    // ---- per-point fold, weights and sub-bin (one point per lane)
    const int j = ch * 64 + lane;
    bool valid = j < sd.z;
    int sbin = 16, offs = 0;
    float wx[NS], cwr[NS], cwi[NS];
    if (valid) {
      const int p = a.perm[sd.y + j];
      valid = (unsigned)p < (unsigned)a.M;
      if (valid) {
        const cplx<float> cv = cT[p];
        const Folded fx = fold_coord((double)static_cast<const CT*>(a.coord[0])[p], a.nf[0]);
        const Folded fy = fold_coord((double)static_cast<const CT*>(a.coord[1])[p], a.nf[1]);
        const int i0x = support_start(fx, NS / 2), i0y = support_start(fy, NS / 2);
        const float tx = (float)((double)i0x - fx.xs), ty = (float)((double)i0y - fy.xs);
        const int sx = (fx.cell - bin_x0) / SB, sy = (fy.cell - bin_y0) / SB;
        const int ox = (i0x - org[0]) - sx * SB, oy = (i0y - org[1]) - sy * SB;  // in [0, SB]
        sbin = sy * 4 + sx;
        offs = ox | (oy << 8);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          wx[k] = es_phi(tx + (float)k, a.es.beta_log2e, a.es.c);
          const float wy = es_phi(ty + (float)k, a.es.beta_log2e, a.es.c);
          cwr[k] = cv.re * wy;
          cwi[k] = cv.im * wy;
        }
      }
    }
    // ---- order the chunk by sub-bin (16 ballots), stage records
    int cnt_s[16], start_s[16];
    int pos = 0, run = 0;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const uint64_t m = __ballot(valid && sbin == s);
      cnt_s[s] = (int)__popcll(m);
      start_s[s] = run;
      if (valid && sbin == s) pos = run + popc_below(m);
      run += cnt_s[s];
    }
    if (valid) {
      float* r = stage + pos * REC;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        r[k] = wx[k];
        r[NS + 2 * k] = cwr[k];
        r[NS + 2 * k + 1] = cwi[k];
      }
      r[3 * NS] = __int_as_float(offs);
    }
    // ---- 4 points per MFMA into the sub-bin accumulators
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int n = cnt_s[s];
      for (int g = 0; g < n; g += 4) {
        float are = 0.0f, aim = 0.0f, b = 0.0f;
        if (g + kk < n) {
          const float* r = stage + (start_s[s] + g + kk) * REC;
          const int o = __float_as_int(r[3 * NS]);
          const int iy = rr - (o >> 8), ix = rr - (o & 0xff);
          if ((unsigned)iy < (unsigned)NS) {
            are = r[NS + 2 * iy];
            aim = r[NS + 2 * iy + 1];
          }
          if ((unsigned)ix < (unsigned)NS) b = r[ix];
        }
        acc_re[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(are, b, acc_re[s], 0, 0, 0);
        acc_im[s] = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, acc_im[s], 0, 0, 0);
      }
    }
  }
  // ---- accumulators -> this wave's tile (windows overlap; one wave, in order)
  const int col = lane & 15, row0 = 4 * (lane >> 4);
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int x = (s & 3) * SB + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = (s >> 2) * SB + row0 + r;
      tre[y * T + x] += acc_re[s][r];
      tim[y * T + x] += acc_im[s][r];
    }
  }
  __syncthreads();
  // ---- sum the wave tiles, flush to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = threadIdx.x; cell < TC; cell += blockDim.x) {
    float vr = 0.0f, vi = 0.0f;
#pragma unroll
    for (int w = 0; w < kMfmaWaves; ++w) {
      vr += smem_f[w * G::WAVE_FLOATS + cell];
      vi += smem_f[w * G::WAVE_FLOATS + TC + cell];
    }
    if (vr == 0.0f && vi == 0.0f) continue;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// Whether a geometry matches the MFMA kernel's bin layout.
inline bool mfma2d_applicable(const Geometry& g, int prec) {
  if (g.dim != 2 || prec != kF32 || g.ns > 8) return false;
  const int bs = 4 * (16 - g.ns);
  return g.bs[0] == bs && g.bs[1] == bs;
}

}  // namespace nufftx
