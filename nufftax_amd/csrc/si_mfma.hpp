// 2D FP32 spreading and interpolation on the matrix cores (gfx950 MFMA).
//
// The contribution of points j to a tile is a sum of outer products,
//   tile[y][x] += sum_j (c_j wy_j[y - y0_j]) * wx_j[x - x0_j]      (nufftax/core/spread.py:465-489)
// i.e. a product A B of a (rows x points) and a (points x cols) matrix whose
// columns/rows carry the ns nonzero kernel weights of each point.  A bin of
// 4 x 4 sub-bins of SB = 16 - ns cells is handled in 16 x 16 windows: every point
// of a sub-bin has its ns x ns support inside the sub-bin's window, so one
// v_mfma_f32_16x16x4f32 accumulates 4 points into a window (spreading) or
// contracts a window with the x-weights of 16 points (interpolation).
//
// Both kernels read the setpts-time point records (sort.hip, prep.hip): each
// subproblem's points in sub-bin order with their tile offsets and t = i0 - xs, so
// no point is refolded here.  A block of 4 waves takes one subproblem; its groups (4 or 16
// points of one sub-bin) are split into 4 contiguous ranges, one per wave, so the
// waves stay balanced whatever the distribution of points over sub-bins.
// Spreading keeps one window accumulator in registers and adds it into the block's
// LDS tile (64-bit fixed point, integer atomics) whenever the sub-bin changes; the
// tile is flushed to the fine grid with global atomics (non-zero cells only).  Interpolation reads the windows from the
// bin's tile, loaded once into LDS.
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

constexpr int kMfmaThreads = 256;  // 4 waves per subproblem block

template <int NS>
struct MfmaGeom {
  static constexpr int SB = 16 - NS;       // sub-bin side (cells)
  static constexpr int BS = 4 * SB;        // bin side: 4 x 4 sub-bins
  static constexpr int T = BS + NS;        // tile side (bin + halo)
  static constexpr int TC = T * T;
  static constexpr int TP = T + 1;         // odd LDS row stride of the interpolation tile
  static constexpr int SREC = 3 * NS + 1;  // spread stage per point: wx[NS], (c wy)[NS] complex, meta
  static constexpr int IREC = 18;          // interp stage per point: wx[8], wy[8], meta, point index
  static constexpr int WAVES = kMfmaThreads / 64;
  // spreading: the block's 64-bit fixed-point tile (re, im; row stride TP) and a 64-point stage per wave
  static constexpr size_t spread_lds() {
    return 2 * (size_t)T * TP * sizeof(unsigned long long) + (size_t)WAVES * 64 * SREC * sizeof(float);
  }
  // interpolation: the tile plus a 64-point stage per wave
  static constexpr size_t interp_lds() { return (2 * (size_t)T * TP + (size_t)WAVES * 64 * IREC) * sizeof(float); }
};

// Orders one wave's LDS writes before its later reads of other lanes' data (and
// those reads before the next writes).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Groups of GP points of one subproblem (setpts-time records in sub-bin order:
// sub-bin s holds records [off[s], off[s+1])); a group never straddles a sub-bin.
// The block's groups are split into 4 contiguous equal ranges, one per wave.
template <int GP>
struct GroupWalk {
  const int32_t* off;  // 17 sub-bin starts (LDS copy)
  int g0, g1;          // this wave's groups
  int s, gs, ge;       // walk state: group range [gs, ge) of sub-bin s
  __device__ __forceinline__ void init(const int32_t* so, int wave) {
    off = so;
    int total = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) total += (so[t + 1] - so[t] + GP - 1) / GP;
    g0 = (int)(((int64_t)total * wave) / 4);
    g1 = (int)(((int64_t)total * (wave + 1)) / 4);
    s = -1;
    gs = ge = 0;
  }
  // advance the walk to the sub-bin holding group g (g non-decreasing, g < total)
  __device__ __forceinline__ void seek(int g) {
    while (g >= ge) {
      ++s;
      gs = ge;
      ge = gs + (off[s + 1] - off[s] + GP - 1) / GP;
    }
  }
  // for any lane: sub-bin and first record of group g, by a scan of the 16 sub-bins
  __device__ __forceinline__ void locate(int g, int& sb, int& first, int& count) const {
    int base = 0;
    sb = 15;
    first = off[16];
    count = 0;
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int n = off[t + 1] - off[t];
      const int ng = (n + GP - 1) / GP;
      if (g >= base && g < base + ng) {
        sb = t;
        first = off[t] + GP * (g - base);
        count = min(GP, n - GP * (g - base));
      }
      base += ng;
    }
  }
};

// ---------------------------------------------------------------- spreading
// A wave walks its groups (4 points of one sub-bin) in batches of 16: lane
// 4 gl + k loads point k of group gl (coalesced records), evaluates its 2 ns
// weights once and stages wx and c wy in the wave's LDS stage; then each group is
// one pair of MFMAs (re, im) into the window accumulator, lane (i, k) feeding
// A[i][k] = c_k wy_k[i - oy_k] and B[k][i] = wx_k[i - ox_k].  Whenever the sub-bin
// changes the window is added into the block's tile in exact 64-bit fixed point
// (integer LDS atomics; scale 2^(50 - e) with 2 n max|c| < 2^e, max|c| over the
// transform from one coalesced reduction), so the tile costs 16 B per cell and no
// wave-private copies: occupancy is what hides this kernel's latencies.
template <int NS>
__global__ void __launch_bounds__(kMfmaThreads) k_spread_mfma2d(SIArgs<float> a, const float4* __restrict__ recs,
                                                                 const int32_t* __restrict__ sub_off,
                                                                 const float* __restrict__ cmax,
                                                                 const cplx<float>* __restrict__ c,
                                                                 cplx<float>* __restrict__ fw) {
  using G = MfmaGeom<NS>;
  using Word = unsigned long long;
  constexpr int SB = G::SB, T = G::T, TP = G::TP, SREC = G::SREC;
  constexpr int TCP = T * TP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Word* sre = reinterpret_cast<Word*>(smem_raw);
  Word* sim = sre + TCP;
  __shared__ int32_t soff[17];
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = uniform(tid >> 6);
  float* stage = reinterpret_cast<float*>(sim + TCP) + wave * 64 * SREC;
  for (int e = tid; e < TCP; e += kMfmaThreads) {
    sre[e] = 0;
    sim[e] = 0;
  }
  if (tid <= 16) soff[tid] = sub_off[(int64_t)sub * 17 + tid];
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const cplx<float>* cT = c + (int64_t)tr * a.M;
  const float4* rT = recs + sd.y;
  __syncthreads();
  // ---- fixed-point scale: every partial sum is below sum_j (|re| + |im|) <= 2 n max|c|
  const double B = 2.0 * (double)sd.z * (double)cmax[tr];
  const bool finite = isfinite(B);
  double scale = 1.0, inv_scale = 1.0;
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);  // B < 2^e
    scale = ldexp(1.0, 50 - e);
    inv_scale = ldexp(1.0, e - 50);
  }
  GroupWalk<4> gw;
  gw.init(soff, wave);

  const int i = lane & 15, k = lane >> 4;
  int cur = -1;
  mfma_f4 acc_re = {0.f, 0.f, 0.f, 0.f}, acc_im = {0.f, 0.f, 0.f, 0.f};
  auto flush = [&](int sbin) {
    const int x = (sbin & 3) * SB + i;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int y = (sbin >> 2) * SB + 4 * k + rr;
      if (acc_re[rr] != 0.0f) atomicAdd(&sre[y * TP + x], TileAcc<float>::quant((double)acc_re[rr], scale));
      if (acc_im[rr] != 0.0f) atomicAdd(&sim[y * TP + x], TileAcc<float>::quant((double)acc_im[rr], scale));
    }
  };
  for (int gb = gw.g0; gb < gw.g1; gb += 16) {
    const int nb = min(16, gw.g1 - gb);
    // ---- stage: lane 4 gl + k takes point k of group gb + gl
    {
      const int gl = lane >> 2, kq = lane & 3;
      int sb, first, count;
      gw.locate(gb + gl, sb, first, count);
      float* r = stage + lane * SREC;
      int meta = -1;
      if (gl < nb && kq < count) {
        const float4 rc = rT[first + kq];
        meta = __float_as_int(rc.z);
        const cplx<float> cv = cT[__float_as_int(rc.w)];
#pragma unroll
        for (int m = 0; m < NS; ++m) {
          r[m] = es_phi(rc.x + (float)m, a.es.beta_log2e, a.es.c);
          const float wy = es_phi(rc.y + (float)m, a.es.beta_log2e, a.es.c);
          r[NS + 2 * m] = cv.re * wy;
          r[NS + 2 * m + 1] = cv.im * wy;
        }
      }
      r[3 * NS] = __int_as_float(meta);
    }
    wave_sync();
    // ---- one MFMA pair per group
    for (int gl = 0; gl < nb; ++gl) {
      gw.seek(gb + gl);
      if (gw.s != cur) {
        if (cur >= 0) flush(cur);
        acc_re = mfma_f4{0.f, 0.f, 0.f, 0.f};
        acc_im = mfma_f4{0.f, 0.f, 0.f, 0.f};
        cur = gw.s;
      }
      const float* r = stage + (4 * gl + k) * SREC;
      const int meta = __float_as_int(r[3 * NS]);
      float are = 0.0f, aim = 0.0f, b = 0.0f;
      if (meta >= 0) {
        const int sbm = (meta >> 24) & 15;
        const int iy = i - (((meta >> 12) & 4095) - (sbm >> 2) * SB), ix = i - ((meta & 4095) - (sbm & 3) * SB);
        if ((unsigned)iy < (unsigned)NS) {
          are = r[NS + 2 * iy];
          aim = r[NS + 2 * iy + 1];
        }
        if ((unsigned)ix < (unsigned)NS) b = r[ix];
      }
      acc_re = __builtin_amdgcn_mfma_f32_16x16x4f32(are, b, acc_re, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, acc_im, 0, 0, 0);
    }
    wave_sync();
  }
  if (cur >= 0) flush(cur);
  __syncthreads();
  // ---- flush non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < T * T; cell += kMfmaThreads) {
    const int l = (cell / T) * TP + cell % T;
    const long long ir = (long long)sre[l], ii = (long long)sim[l];
    if (finite && ir == 0 && ii == 0) continue;
    const float vr = finite ? (float)((double)ir * inv_scale) : NAN;
    const float vi = finite ? (float)((double)ii * inv_scale) : NAN;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// cmax[t] = max over j of max(|re c[t][j]|, |im c[t][j]|), as float bits (a NaN or an
// infinity propagates as the largest pattern): per-block maxima into part[t][block],
// then one block per transform reduces them (at most kAbsmaxBlocks partials).  No atomics.

static __global__ void __attribute__((unused)) __launch_bounds__(256) k_absmax_part(const cplx<float>* __restrict__ c,
                                                                                   int64_t M, unsigned* __restrict__ part) {
  __shared__ unsigned wmax[4];
  const cplx<float>* cT = c + (int64_t)blockIdx.y * M;
  unsigned m = 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; j + 3 * stride < M; j += 4 * stride) {
    cplx<float> v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = cT[j + u * stride];
#pragma unroll
    for (int u = 0; u < 4; ++u) m = max(m, max(__float_as_uint(fabsf(v[u].re)), __float_as_uint(fabsf(v[u].im))));
  }
  for (; j < M; j += stride) {
    const cplx<float> v = cT[j];
    m = max(m, max(__float_as_uint(fabsf(v.re)), __float_as_uint(fabsf(v.im))));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3]));
}

static __global__ void __attribute__((unused)) __launch_bounds__(256) k_absmax_final(const unsigned* __restrict__ part,
                                                                                    int nparts, float* __restrict__ cmax) {
  __shared__ unsigned wmax[4];
  unsigned m = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x) m = max(m, part[(int64_t)blockIdx.x * nparts + b]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off, 64));
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) cmax[blockIdx.x] = __uint_as_float(max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}

// ---------------------------------------------------------------- interpolation
// c_j = sum_y wy_j[y] sum_x T[y][x] wx_j[x]                  (nufftax/core/spread.py:497-540)
// One group = 16 points of one sub-bin: D = T_win Wx, a (16 rows x 16 cols) x
// (16 cols x 16 points) product in 4-column steps of v_mfma_f32_16x16x4f32 (A =
// window columns from the LDS tile, B = x-weights), then each lane folds its 4
// rows of D with the y-weights and 2 cross-lane adds finish the sum over y.
// A wave walks its groups in batches of 4: lane 16 gl + j loads point j of group
// gl, evaluates its 2 ns weights once and stages them in the wave's LDS stage.
template <int NS>
__global__ void __launch_bounds__(kMfmaThreads) k_interp_mfma2d(SIArgs<float> a, const float4* __restrict__ recs,
                                                                 const int32_t* __restrict__ sub_off,
                                                                 const cplx<float>* __restrict__ fw,
                                                                 cplx<float>* __restrict__ out) {
  using G = MfmaGeom<NS>;
  constexpr int SB = G::SB, T = G::T, TP = G::TP, IREC = G::IREC;
  extern __shared__ __align__(16) float smem_f[];
  float* tre = smem_f;
  float* tim = tre + T * TP;
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = uniform(tid >> 6);
  float* stage = tim + T * TP + wave * 64 * IREC;
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < T * T; cell += kMfmaThreads) {
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    const cplx<float> v = fT[gi];
    const int l = (cell / T) * TP + cell % T;
    tre[l] = v.re;
    tim[l] = v.im;
  }
  __syncthreads();
  cplx<float>* oT = out + (int64_t)tr * a.M;
  const float4* rT = recs + sd.y;
  __shared__ int32_t soff[17];
  if (tid <= 16) soff[tid] = sub_off[(int64_t)sub * 17 + tid];
  __syncthreads();
  GroupWalk<16> gw;
  gw.init(soff, wave);
  const int j = lane & 15, q = lane >> 4;
  for (int gb = gw.g0; gb < gw.g1; gb += 4) {
    const int nb = min(4, gw.g1 - gb);
    // ---- stage: lane 16 gl + jj takes point jj of group gb + gl
    {
      const int gl = lane >> 4, jj = lane & 15;
      int sb, first, count;
      gw.locate(gb + gl, sb, first, count);
      float* r = stage + lane * IREC;
      int meta = -1, p = 0;
      float4 rc = make_float4(0.f, 0.f, 0.f, 0.f);
      if (gl < nb && jj < count) {
        rc = rT[first + jj];
        meta = __float_as_int(rc.z);
        p = __float_as_int(rc.w);
      }
#pragma unroll
      for (int m = 0; m < 8; ++m) {
        r[m] = (meta >= 0 && m < NS) ? es_phi(rc.x + (float)m, a.es.beta_log2e, a.es.c) : 0.0f;
        r[8 + m] = (meta >= 0 && m < NS) ? es_phi(rc.y + (float)m, a.es.beta_log2e, a.es.c) : 0.0f;
      }
      r[16] = __int_as_float(meta);
      r[17] = __int_as_float(p);
    }
    wave_sync();
    for (int gl = 0; gl < nb; ++gl) {
      gw.seek(gb + gl);
      const int s = gw.s;
      const float* r = stage + (16 * gl + j) * IREC;
      const int meta = __float_as_int(r[16]);
      const int ox = (meta & 4095) - (s & 3) * SB, oy = ((meta >> 12) & 4095) - (s >> 2) * SB;
      const int rowA = ((s >> 2) * SB + j) * TP + (s & 3) * SB + q;
      mfma_f4 dre = {0.f, 0.f, 0.f, 0.f}, dim = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int xc = 0; xc < 4; ++xc) {
        const float are = tre[rowA + 4 * xc], aim = tim[rowA + 4 * xc];
        const int kk = 4 * xc + q - ox;
        const float b = (meta >= 0 && (unsigned)kk < (unsigned)NS) ? r[kk] : 0.0f;
        dre = __builtin_amdgcn_mfma_f32_16x16x4f32(are, b, dre, 0, 0, 0);
        dim = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, b, dim, 0, 0, 0);
      }
      float pre = 0.0f, pim = 0.0f;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int ky = 4 * q + e - oy;
        const float wy = (unsigned)ky < (unsigned)NS ? r[8 + ky] : 0.0f;
        pre = fmaf(wy, dre[e], pre);
        pim = fmaf(wy, dim[e], pim);
      }
      pre += __shfl_xor(pre, 16, 64);
      pim += __shfl_xor(pim, 16, 64);
      pre += __shfl_xor(pre, 32, 64);
      pim += __shfl_xor(pim, 32, 64);
      if (q == 0 && meta >= 0) oT[__float_as_int(r[17])] = cplx<float>{pre, pim};
    }
    wave_sync();
  }
}

// The matrix-core path (setpts-time point records + matrix-core spread/interp) is
// opt-in (NUFFTX_MFMA=1): on randomly ordered points the records cost the bin-sort
// more than they save the spread/interp kernels, so the LDS-tiled kernels are the
// default.  NUFFTX_NO_MFMA=1 forces it off.
inline bool use_mfma() { return std::getenv("NUFFTX_MFMA") != nullptr && std::getenv("NUFFTX_NO_MFMA") == nullptr; }

// Whether a geometry matches the MFMA kernel's bin layout.
inline bool mfma2d_applicable(const Geometry& g, int prec) {
  // ns <= 8 keeps SB = 16 - ns >= 8, which the spread kernel's phased tile update needs
  if (g.dim != 2 || prec != kF32 || g.ns > 8) return false;
  const int bs = 4 * (16 - g.ns);
  return g.bs[0] == bs && g.bs[1] == bs;
}

}  // namespace nufftx
