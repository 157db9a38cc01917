// 2D FP32 spreading and interpolation on the matrix cores (gfx950 MFMA).
//
// The contribution of points j to a tile is a sum of outer products,
//   tile[y][x] += sum_j (c_j wy_j[y - y0_j]) * wx_j[x - x0_j]      (nufftax/core/spread.py:465-489)
// i.e. a product A B of a (rows x points) and a (points x cols) matrix whose
// columns/rows carry the ns nonzero kernel weights of each point.  A bin of
// 4 x 4 sub-bins of SB = 16 - ns cells is spread with 16x16 windows: every point
// of a sub-bin has its ns x ns support inside the sub-bin's 16 x 16 window, so one
// v_mfma_f32_16x16x4f32 accumulates 4 points into the window (4 x 16 A operand =
// c * y-weights, 16 x 4 B operand = x-weights).
//
// Spreading: a block of 4 waves takes one subproblem (<= maxsub points of one
// bin) in segments of kSpreadSeg points.  Each thread folds one point (FP64),
// evaluates its 2 ns weights once and the block counting-sorts the segment by
// sub-bin into LDS records.  Wave w owns one sub-bin window of each of the 4
// parity classes (sx & 1, sy & 1) and keeps their accumulators in registers for
// the whole subproblem; at the end the 4 classes are added into the LDS tile one
// after another (windows of a class are disjoint), and non-zero tile cells are
// flushed to the fine grid with global atomics.  Nothing is accumulated with LDS
// atomics.
//
// MFMA operand layout (v_mfma_f32_16x16x4f32, 64 lanes):
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k,
//   D[i][j] in lane j + 16 (i / 4), element i % 4.
#pragma once

#include <cstdlib>

#include "si_kernels.hpp"

namespace nufftx {

typedef float mfma_f4 __attribute__((ext_vector_type(4)));

// A value the caller knows to be the same in every lane of the wave, moved to a
// scalar register so that loops bounded by it are scalar (non-divergent) loops.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

constexpr int kSpreadSeg = 256;      // points staged per segment (one per thread)
constexpr int kSpreadThreads = 256;  // 4 waves

template <int NS>
struct MfmaGeom {
  static constexpr int SB = 16 - NS;       // sub-bin side (cells)
  static constexpr int NSUB = 4;           // sub-bins per bin side
  static constexpr int BS = NSUB * SB;     // bin side
  static constexpr int T = BS + NS;        // tile side
  static constexpr int TC = T * T;
  static constexpr int REC = 3 * NS + 1;   // staged floats per point: wx[NS], (c wy)[NS] complex, offsets
  static constexpr size_t lds_bytes() {
    return (2 * (size_t)TC + (size_t)kSpreadSeg * REC) * sizeof(float);
  }
};

// Sub-bin s = sy * 4 + sx belongs to parity class (sx & 1) | (sy & 1) << 1; the 4
// windows of a class are disjoint (2 SB >= 16).  Wave w owns the w-th sub-bin of
// every class.
__device__ __forceinline__ int owned_subbin(int wave, int cls) {
  const int sx = (cls & 1) + 2 * (wave & 1), sy = (cls >> 1) + 2 * (wave >> 1);
  return sy * 4 + sx;
}

template <typename CT, int NS>
__global__ void __launch_bounds__(kSpreadThreads) k_spread_mfma2d(SIArgs<float> a, const cplx<float>* __restrict__ c,
                                                                   cplx<float>* __restrict__ fw) {
  using G = MfmaGeom<NS>;
  constexpr int SB = G::SB, T = G::T, TC = G::TC, REC = G::REC;
  extern __shared__ __align__(16) float smem_f[];
  __shared__ int cnt[17];
  float* tre = smem_f;
  float* tim = tre + TC;
  float* rec = tim + TC;

  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = uniform(tid >> 6);
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const int bin_x0 = org[0] + NS / 2, bin_y0 = org[1] + NS / 2;
  for (int i = tid; i < TC; i += kSpreadThreads) {
    tre[i] = 0.0f;
    tim[i] = 0.0f;
  }
  const cplx<float>* cT = c + (int64_t)tr * a.M;

  mfma_f4 acc_re[4], acc_im[4];  // this wave's window of each parity class
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    acc_re[k] = mfma_f4{0.f, 0.f, 0.f, 0.f};
    acc_im[k] = mfma_f4{0.f, 0.f, 0.f, 0.f};
  }
  const int kk = lane >> 4;  // point of the 4-point group this lane feeds
  const int rr = lane & 15;  // window row (A) / column (B) this lane feeds

  for (int seg0 = 0; seg0 < sd.z; seg0 += kSpreadSeg) {
    if (tid <= 16) cnt[tid] = 0;
    __syncthreads();
    // ---- one point per thread: fold (FP64), weights, sub-bin, slot
    const int j = seg0 + tid;
    int s = -1, slot = 0, offs = 0;
    float wx[NS], cwr[NS], cwi[NS];
    if (j < sd.z) {
      const int p = a.perm[sd.y + j];
      if ((unsigned)p < (unsigned)a.M) {
        const cplx<float> cv = cT[p];
        const Folded fx = fold_coord((double)static_cast<const CT*>(a.coord[0])[p], a.nf[0]);
        const Folded fy = fold_coord((double)static_cast<const CT*>(a.coord[1])[p], a.nf[1]);
        const int i0x = support_start(fx, NS / 2), i0y = support_start(fy, NS / 2);
        const float tx = (float)((double)i0x - fx.xs), ty = (float)((double)i0y - fy.xs);
        const int sx = (fx.cell - bin_x0) / SB, sy = (fy.cell - bin_y0) / SB;
        const int ox = (i0x - org[0]) - sx * SB, oy = (i0y - org[1]) - sy * SB;  // in [0, SB]
        s = sy * 4 + sx;
        offs = ox | (oy << 8);
#pragma unroll
        for (int k = 0; k < NS; ++k) {
          wx[k] = es_phi(tx + (float)k, a.es.beta_log2e, a.es.c);
          const float wy = es_phi(ty + (float)k, a.es.beta_log2e, a.es.c);
          cwr[k] = cv.re * wy;
          cwi[k] = cv.im * wy;
        }
        slot = atomicAdd(&cnt[s], 1);
      }
    }
    __syncthreads();
    // ---- exclusive scan of the 16 sub-bin counts (one wave)
    if (wave == 0) {
      const int v = lane < 16 ? cnt[lane] : 0;
      int incl = v;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      if (lane < 16) cnt[lane] = incl - v;  // start of sub-bin lane
      if (lane == 15) cnt[16] = incl;
    }
    __syncthreads();
    // ---- stage records in sub-bin order
    if (s >= 0) {
      float* r = rec + (cnt[s] + slot) * REC;
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        r[k] = wx[k];
        r[NS + 2 * k] = cwr[k];
        r[NS + 2 * k + 1] = cwi[k];
      }
      r[3 * NS] = __int_as_float(offs);
    }
    __syncthreads();
    // ---- this wave's 4 sub-bins: 4 points per MFMA into the window accumulators
#pragma unroll
    for (int cls = 0; cls < 4; ++cls) {
      const int sb = owned_subbin(wave, cls);
      const int st = uniform(cnt[sb]), n = uniform(cnt[sb + 1]) - st;
      for (int g = 0; g < n; g += 4) {
        float are = 0.0f, aim = 0.0f, b = 0.0f;
        if (g + kk < n) {
          const float* r = rec + (st + g + kk) * REC;
          const int o = __float_as_int(r[3 * NS]);
          const int iy = rr - (o >> 8), ix = rr - (o & 0xff);
          if ((unsigned)iy < (unsigned)NS) {
            are = r[NS + 2 * iy];
            aim = r[NS + 2 * iy + 1];
          }
          if ((unsigned)ix < (unsigned)NS) b = r[ix];
        }
        acc_re[cls] = __builtin_amdgcn_mfma_f32_16x16x4f32(are, b, acc_re[cls], 0, 0, 0);
        acc_im[cls] = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, acc_im[cls], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // ---- windows -> tile, one parity class at a time (disjoint windows, one per wave)
  const int col = lane & 15, row0 = 4 * (lane >> 4);
#pragma unroll
  for (int cls = 0; cls < 4; ++cls) {
    const int sb = owned_subbin(wave, cls);
    const int x = (sb & 3) * SB + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int y = (sb >> 2) * SB + row0 + r;
      tre[y * T + x] += acc_re[cls][r];
      tim[y * T + x] += acc_im[cls][r];
    }
    __syncthreads();
  }
  // ---- flush non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < TC; cell += kSpreadThreads) {
    const float vr = tre[cell], vi = tim[cell];
    if (vr == 0.0f && vi == 0.0f) continue;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// 2D FP32 interpolation on the matrix cores, same bin / sub-bin geometry.
//
// c_j = sum_y wy_j[y] sum_x T[y][x] wx_j[x]                  (nufftax/core/spread.py:497-540)
// For 16 points of one sub-bin the inner sum is a (16 rows x 16 cols) x (16 cols x
// 16 points) product, D = T_win Wx, done as 4-column steps of v_mfma_f32_16x16x4f32
// (A = window columns, B = x-weights of the 16 points); each lane then folds its 4
// rows of D with the y-weights and 2 cross-lane adds finish the sum over y.
//
// A block takes one subproblem: the bin's tile (bin + halo) is loaded into LDS
// once.  Points are taken in segments of kInterpSeg: each is folded (FP64), its
// 2 ns weights are evaluated once, and the records are counting-sorted in LDS by
// (sub-bin, x offset), so every group of 16 points shares a window and a narrow
// range of x offsets: column steps outside the group's support are skipped.
constexpr int kInterpSeg = 512;
constexpr int kInterpThreads = 256;

template <int NS>
struct InterpGeom {
  static constexpr int SB = 16 - NS;
  static constexpr int T = 4 * SB + NS;
  static constexpr int TP = T + 1;          // odd LDS row stride (bank spread of the A reads)
  static constexpr int TCP = T * TP;
  static constexpr int NKEY = 16 * (SB + 1);
  static constexpr int MAXG = kInterpSeg / 16 + 16;
  static constexpr int REC = 2 * NS + 2;    // wx[NS], wy[NS], offsets | sub-bin, point index
  static constexpr size_t lds_bytes() {
    return 2 * (size_t)TCP * sizeof(float) + (size_t)kInterpSeg * REC * sizeof(float);
  }
};

template <typename CT, int NS>
__global__ void __launch_bounds__(kInterpThreads) k_interp_mfma2d(SIArgs<float> a, const cplx<float>* __restrict__ fw,
                                                                   cplx<float>* __restrict__ out) {
  using G = InterpGeom<NS>;
  constexpr int SB = G::SB, T = G::T, TP = G::TP, TCP = G::TCP, NKEY = G::NKEY, REC = G::REC;
  constexpr int PPT = kInterpSeg / kInterpThreads;
  extern __shared__ __align__(16) float smem_f[];
  __shared__ int cnt[NKEY + 1];
  __shared__ int2 grp[G::MAXG];
  __shared__ int ngrp;
  float* tre = smem_f;
  float* tim = tre + TCP;
  float* rec = tim + TCP;

  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = uniform(tid >> 6);
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const int bin_x0 = org[0] + NS / 2, bin_y0 = org[1] + NS / 2;

  // ---- tile (bin + halo) -> LDS
  const cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < T * T; cell += kInterpThreads) {
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    const cplx<float> v = fT[gi];
    const int l = (cell / T) * TP + cell % T;
    tre[l] = v.re;
    tim[l] = v.im;
  }
  cplx<float>* oT = out + (int64_t)tr * a.M;
  const int j = lane & 15, q = lane >> 4;

  for (int seg0 = 0; seg0 < sd.z; seg0 += kInterpSeg) {
    for (int k = tid; k <= NKEY; k += kInterpThreads) cnt[k] = 0;
    __syncthreads();
    // ---- fold, weights, sub-bin, key; per-key slot by LDS atomics
    float w[PPT][2 * NS];
    int meta[PPT], pidx[PPT], key[PPT], slot[PPT];
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      const int jj = seg0 + u * kInterpThreads + tid;
      key[u] = -1;
      if (jj < sd.z) {
        const int p = a.perm[sd.y + jj];
        if ((unsigned)p < (unsigned)a.M) {
          const Folded fx = fold_coord((double)static_cast<const CT*>(a.coord[0])[p], a.nf[0]);
          const Folded fy = fold_coord((double)static_cast<const CT*>(a.coord[1])[p], a.nf[1]);
          const int i0x = support_start(fx, NS / 2), i0y = support_start(fy, NS / 2);
          const int sx = (fx.cell - bin_x0) / SB, sy = (fy.cell - bin_y0) / SB;
          const int ox = (i0x - org[0]) - sx * SB, oy = (i0y - org[1]) - sy * SB;
          const int s = sy * 4 + sx;
          const float tx = (float)((double)i0x - fx.xs), ty = (float)((double)i0y - fy.xs);
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            w[u][k] = es_phi(tx + (float)k, a.es.beta_log2e, a.es.c);
            w[u][NS + k] = es_phi(ty + (float)k, a.es.beta_log2e, a.es.c);
          }
          meta[u] = ox | (oy << 8) | (s << 16);
          pidx[u] = p;
          key[u] = s * (SB + 1) + ox;
          slot[u] = atomicAdd(&cnt[key[u]], 1);
        }
      }
    }
    __syncthreads();
    // ---- exclusive scan of the key counts (wave 0), then the group list (16 lanes)
    if (wave == 0) {
      constexpr int KPL = (NKEY + 63) / 64;
      int v[KPL], sum = 0;
#pragma unroll
      for (int t = 0; t < KPL; ++t) {
        const int k = lane * KPL + t;
        v[t] = k < NKEY ? cnt[k] : 0;
        sum += v[t];
      }
      int incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      int run = incl - sum;
#pragma unroll
      for (int t = 0; t < KPL; ++t) {
        const int k = lane * KPL + t;
        if (k < NKEY) cnt[k] = run;
        run += v[t];
      }
      if (lane == 63) cnt[NKEY] = incl;
    }
    __syncthreads();
    if (wave == 0) {
      int st = 0, n = 0, ng = 0;
      if (lane < 16) {
        st = cnt[lane * (SB + 1)];
        n = cnt[(lane + 1) * (SB + 1)] - st;
        ng = (n + 15) >> 4;
      }
      int incl = ng;
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        const int o = __shfl_up(incl, off, 64);
        if (lane >= off) incl += o;
      }
      if (lane < 16) {
        const int g0 = incl - ng;
        for (int g = 0; g < ng; ++g) grp[g0 + g] = make_int2(st + 16 * g, lane | (min(16, n - 16 * g) << 8));
      }
      if (lane == 15) ngrp = incl;
    }
    // ---- scatter records in key order
#pragma unroll
    for (int u = 0; u < PPT; ++u) {
      if (key[u] < 0) continue;
      float* r = rec + (cnt[key[u]] + slot[u]) * REC;
#pragma unroll
      for (int k = 0; k < 2 * NS; ++k) r[k] = w[u][k];
      r[2 * NS] = __int_as_float(meta[u]);
      r[2 * NS + 1] = __int_as_float(pidx[u]);
    }
    __syncthreads();
    // ---- groups of 16 points: D = T_win Wx on the matrix cores, then the y sum
    const int ng_total = uniform(ngrp);
    for (int g = wave; g < ng_total; g += kInterpThreads / 64) {
      const int2 gd = grp[g];
      const int gx = uniform(gd.x), gy = uniform(gd.y);
      const int s = gy & 255, gn = gy >> 8;
      const bool valid = j < gn;
      const float* r = rec + (gx + min(j, gn - 1)) * REC;
      const int m = __float_as_int(r[2 * NS]);
      const int ox = m & 255, oy = (m >> 8) & 255;
      const int ox_lo = __builtin_amdgcn_readlane(ox, 0), ox_hi = __builtin_amdgcn_readlane(ox, gn - 1);
      const int xc_lo = ox_lo >> 2, xc_hi = min(3, (ox_hi + NS - 1) >> 2);
      const int rowA = ((s >> 2) * SB + j) * TP + (s & 3) * SB + q;
      mfma_f4 dre = {0.f, 0.f, 0.f, 0.f}, dim = {0.f, 0.f, 0.f, 0.f};
      for (int xc = xc_lo; xc <= xc_hi; ++xc) {
        const float are = tre[rowA + 4 * xc], aim = tim[rowA + 4 * xc];
        const int kk = 4 * xc + q - ox;
        const float b = (valid && (unsigned)kk < (unsigned)NS) ? r[kk] : 0.0f;
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(are, b, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, dim, 0, 0, 0);
      }
      float pre = 0.0f, pim = 0.0f;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ky = 4 * q + e - oy;
        const float wy = (unsigned)ky < (unsigned)NS ? r[NS + ky] : 0.0f;
        pre = fmaf(wy, dre[e], pre);
        pim = fmaf(wy, dim[e], pim);
      }
      pre += __shfl_xor(pre, 16, 64);
      pim += __shfl_xor(pim, 16, 64);
      pre += __shfl_xor(pre, 32, 64);
      pim += __shfl_xor(pim, 32, 64);
      if (q == 0 && valid) oT[__float_as_int(r[2 * NS + 1])] = cplx<float>{pre, pim};
    }
    __syncthreads();
  }
}

// Kernel choice for 2D FP32 ns <= 8: the matrix-core spread is the default (it beats
// the LDS-tiled fixed-point spread), the matrix-core interpolation is opt-in
// (NUFFTX_MFMA=1) until it beats the LDS-tiled gather.  NUFFTX_NO_MFMA=1 forces both off.
inline bool use_mfma_spread() { return std::getenv("NUFFTX_NO_MFMA") == nullptr; }
inline bool use_mfma_interp() {
  return std::getenv("NUFFTX_MFMA") != nullptr && std::getenv("NUFFTX_NO_MFMA") == nullptr;
}

// Whether a geometry matches the MFMA kernel's bin layout.
inline bool mfma2d_applicable(const Geometry& g, int prec) {
  // ns <= 8 keeps SB = 16 - ns >= 8, which the spread kernel's phased tile update needs
  if (g.dim != 2 || prec != kF32 || g.ns > 8) return false;
  const int bs = 4 * (16 - g.ns);
  return g.bs[0] == bs && g.bs[1] == bs;
}

}  // namespace nufftx
