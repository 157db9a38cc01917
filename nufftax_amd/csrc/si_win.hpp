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
// The points come packed in sorted order, {x word, y word, re c, im c} per point and
// transform (pointset_pack2d), word = l << 25 | u: l the support start in the bin's tile,
// u = round((t + ns/2 + 1) 2^24), t = i0 - xs from the FP64 fold.
//
// Block = one subproblem (<= 1024 points of one bin), 8 waves:
//   1. order: every point's support start gives its sub-bin; the points are ranked per
//      sub-bin inside each wave (four ballots) and their indices written into LDS in
//      sub-bin order (stable), each sub-bin padded to a multiple of 4 with
//      zero-strength entries; the block also takes max(|re c|, |im c|);
//   2. accumulate: the groups of 4 are split into equal contiguous ranges, one per
//      wave, walked in batches of 16 points.  Per batch, the wave re-reads the batch's
//      records (cache-resident) and each lane evaluates ns/2 kernel weights of one
//      dimension of one point (ES kernel in registers), written into entries 16 - ns ..
//      15 of an LDS row of 17 floats whose other entries stay zero: the weight of window
//      row (or column) i of a point with window offset o is entry 16 - ns + i - o of its
//      row, read straight on into the next row's zeros (i - o runs over
//      [-(16 - ns), 15]), so the MFMA operands A[i][k] = c_k wy_k(i), B[k][i] = wx_k(i)
//      are two LDS reads and two multiplies per lane, no bounds checks.  When the wave's
//      walk leaves a sub-bin, the window is added into the block's tile in fixed point
//      (integer LDS atomics, 4 rows x 16 cells per step);
//   3. flush: the tile's non-zero cells go to the fine grid with FP32 global atomics.
// Windows accumulate in FP32 (like the reference's segment_sum).  Tiles (PK = false): two
// 64-bit words per cell, scale 2^(50 - e) with 2^e above 2 n max(|re c|, |im c|) over the
// subproblem's n points, so every partial sum stays below 2^50 and the tile sum of the
// windows is exact.  Packed tiles (PK = true, the default): one 64-bit word per cell, the
// real part in its high and the imaginary part in its low 32 bits, scale 2^(30 - e) with
// 2^e above S = sum_j max(|re c_j|, |im c_j|), which bounds every window value and every
// final tile value of either part: the word adds modulo 2^64, so only the final sums must
// fit 31 bits (they do, with a factor 2 to spare) and partial sums may wrap.  A window value
// is then rounded to S 2^-30 (about 2^-20 of the subproblem's strength sum), 4 adds per cell
// and subproblem on average -- near FP32 rounding -- for half the tile and half the
// atomics, which lets four blocks share a CU.
//
// MFMA operand layout (v_mfma_f32_16x16x4f32, 64 lanes):
//   A[i][k] in lane i + 16 k,  B[k][j] in lane j + 16 k,  D[i][j] in lane j + 16 (i / 4), element i % 4.
#pragma once

#include <cstdlib>

#include "si_kernels.hpp"

namespace nufftx {

typedef float win_f4 __attribute__((ext_vector_type(4)));

constexpr int kWinThreads = 512;  // 8 waves per subproblem
constexpr int kWinPpt = 2;        // points per thread while ordering: subproblems of <= 1024 points
constexpr int kWinBatch = 16;     // points per weight batch of a wave
constexpr int kWinRow = 17;       // LDS row stride (floats): odd, so that the rows of a batch fall on distinct banks
constexpr int kWinRowF = kWinBatch * kWinRow + 16;       // floats of one row block (+ trailing zeros)
constexpr int kWinWaveF = 2 * kWinRowF + 4 * kWinBatch;  // floats per wave: x rows, y rows, meta
static_assert(kWinPpt * kWinThreads == 1024, "win2d_geometry admits subproblems of up to 1024 points");

template <int NS>
struct WinGeom {
  static constexpr int SB = 16 - NS;  // sub-bin side (support starts); window 16 = SB + NS
  static constexpr int BS = 4 * SB;   // bin side
  static constexpr int T = BS + NS;   // tile side
};

__host__ __device__ inline size_t win_order_bytes(int maxsub) { return ((size_t)(maxsub + 3 * 16) * sizeof(uint16_t) + 15) / 16 * 16; }

// LDS bytes of the window spread (dynamic part) for subproblems of maxsub points
template <int NS, bool PK>
inline size_t win_spread_lds(int maxsub) {
  constexpr int T = WinGeom<NS>::T;
  return (PK ? 1 : 2) * (size_t)T * T * sizeof(unsigned long long) + win_order_bytes(maxsub) +
         (size_t)(kWinThreads / 64) * kWinWaveF * sizeof(float);
}

// Orders a wave's LDS writes before its later reads of other lanes' data (and those
// reads before the next writes).
__device__ __forceinline__ void wave_sync_win() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int NS, bool PK>
__global__ void __launch_bounds__(kWinThreads) k_spread_win2d(SIArgs<float> a, int maxsub,
                                                              const uint4* __restrict__ packed,
                                                              cplx<float>* __restrict__ fw) {
  using G = WinGeom<NS>;
  using Word = unsigned long long;
  constexpr int SB = G::SB, T = G::T, TC = T * T;
  constexpr int W = kWinThreads / 64;
  constexpr uint16_t kPad = 0xFFFF;  // order entry of a zero-strength padding point
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Word* sre = reinterpret_cast<Word*>(smem_raw);
  Word* sim = sre + TC;  // unused when packed
  uint16_t* order = reinterpret_cast<uint16_t*>(sre + (PK ? 1 : 2) * TC);
  __shared__ int wcnt[kWinPpt * W * 16];  // per (pass, wave, sub-bin): count, then first slot
  __shared__ int cnt[16], start[17];
  __shared__ unsigned smax;
  __shared__ float ssum[W];
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int4 sd = a.subprob[sub];
  const int n = sd.z;
  const uint4* __restrict__ pk = packed + (int64_t)tr * a.M + sd.y;
  float* const lf = reinterpret_cast<float*>(smem_raw);
  const int wof = (int)(((PK ? 1 : 2) * (size_t)TC * sizeof(Word) + win_order_bytes(maxsub)) / 4) + wave * kWinWaveF;
  float* rowx = lf + wof;         // x weights: row q at 17 q, weights at 17 q + 16 - NS .. 17 q + 15
  float* rowy = rowx + kWinRowF;  // y weights, same layout
  int4* meta = reinterpret_cast<int4*>(rowy + kWinRowF);  // {y read index, x read index, re c, im c}
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  for (int e = tid; e < TC; e += kWinThreads) {
    sre[e] = 0;
    if constexpr (!PK) sim[e] = 0;
  }
  for (int e = lane; e < kWinRowF; e += 64)  // the zero parts of the weight rows stay zero
    if (e % kWinRow < 16 - NS || e % kWinRow == 16 || e >= kWinBatch * kWinRow) {
      rowx[e] = 0.0f;
      rowy[e] = 0.0f;
    }
  if (lane < kWinBatch)  // read indices inside the wave's rows until the first batch sets them
    meta[lane] = make_int4(wof + kWinRowF + kWinRow * lane, wof + kWinRow * lane, 0, 0);
  for (int e = tid; e < kWinPpt * W * 16; e += kWinThreads) wcnt[e] = 0;
  if (tid == 0) smax = 0u;
  __syncthreads();

  // ---- 1. order
  int sb[kWinPpt], rank[kWinPpt];
  unsigned mymax = 0;
  float mysum = 0.0f;
  uint4 pr[kWinPpt];
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r) {  // all loads first
    const int j = tid + r * kWinThreads;
    pr[r] = j < n ? pk[j] : make_uint4(0u, 0u, 0u, 0u);
  }
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r) {
    const int j = tid + r * kWinThreads;
    sb[r] = -1;
    if (j >= n) continue;
    const uint4 p = pr[r];
    const int sx = min((int)(p.x >> 25) / SB, 3), sy = min((int)(p.y >> 25) / SB, 3);
    const float are = fabsf(__uint_as_float(p.z)), aim = fabsf(__uint_as_float(p.w));
    mymax = max(mymax, max(__float_as_uint(are), __float_as_uint(aim)));
    mysum += fmaxf(are, aim) + 0.0f * (are + aim);  // NaN in either part propagates
    sb[r] = sy * 4 + sx;
  }
  // rank of each point among the wave's points of its sub-bin (match by four ballots)
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r) {
    const bool valid = sb[r] >= 0;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint64_t bal = __ballot(valid && ((sb[r] >> b) & 1));
      m &= ((sb[r] >> b) & 1) ? bal : ~bal;
    }
    rank[r] = popc_below(m);
    if (valid && rank[r] == 0) wcnt[(r * W + wave) * 16 + sb[r]] = (int)__popcll(m);
  }
  // block maximum of |re c|, |im c| (non-negative floats order as their bit patterns;
  // a NaN pattern is above every finite one)
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    mymax = max(mymax, (unsigned)__shfl_xor((int)mymax, off, 64));
    mysum += __shfl_xor(mysum, off, 64);
  }
  if (lane == 0) {
    atomicMax(&smax, mymax);
    ssum[wave] = mysum;
  }
  __syncthreads();
  if (tid < 16) {  // per sub-bin count
    int c = 0;
    for (int e = 0; e < kWinPpt * W; ++e) c += wcnt[e * 16 + tid];
    cnt[tid] = c;
  }
  __syncthreads();
  if (tid == 0) {  // padded sub-bin starts (multiples of 4)
    int acc = 0;
    for (int s = 0; s < 16; ++s) {
      start[s] = acc;
      acc += (cnt[s] + 3) & ~3;
    }
    start[16] = acc;
  }
  __syncthreads();
  if (tid < 16) {  // first slot of every (pass, wave) run of the sub-bin
    int run = start[tid];
    for (int e = 0; e < kWinPpt * W; ++e) {
      const int v = wcnt[e * 16 + tid];
      wcnt[e * 16 + tid] = run;
      run += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kWinPpt; ++r)
    if (sb[r] >= 0) order[wcnt[(r * W + wave) * 16 + sb[r]] + rank[r]] = (uint16_t)(tid + r * kWinThreads);
  if (tid < 16 * 3) {  // zero-strength padding entries
    const int s = tid / 3, q = tid % 3;
    const int slot = start[s] + cnt[s] + q;
    if (slot < start[s + 1]) order[slot] = kPad;
  }
  const float cm = __uint_as_float(smax);
  double B;
  if constexpr (PK) {
    float S = 0.0f;
#pragma unroll
    for (int w = 0; w < W; ++w) S += ssum[w];
    B = (double)S * (1.0 + 1.0 / 1024.0) + (double)cm;  // above the exact sum despite FP32 rounding
  } else {
    B = 2.0 * (double)n * (double)cm;
  }
  const bool finite = isfinite(B) && isfinite(cm);
  constexpr int kBits = PK ? 30 : 50;
  double scale = 1.0, inv_scale = 1.0;
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);  // B < 2^e
    scale = ldexp(1.0, kBits - e);
    inv_scale = ldexp(1.0, e - kBits);
  }
  __syncthreads();

  // ---- 2. accumulate
  const int NG = start[16] >> 2;
  const int g0 = (int)(((int64_t)NG * wave) / W), g1 = (int)(((int64_t)NG * (wave + 1)) / W);
  const int i = lane & 15, k = lane >> 4;
  // weight staging: lane -> (point sq, dimension sdim, weights sh NS/2 .. sh NS/2 + NS/2 - 1)
  const int sq = lane & (kWinBatch - 1), sdim = (lane >> 4) & 1, sh = lane >> 5;
  const float bl2e = a.es.beta_log2e, kc = a.es.c;
  float* const myrow = (sdim ? rowy : rowx) + kWinRow * sq + (16 - NS) + sh * (NS / 2);
  const int ry0 = wof + kWinRowF + (16 - NS), rx0 = wof + (16 - NS);  // first weight of row 0 (y, x)
  win_f4 acc_re = {0.f, 0.f, 0.f, 0.f}, acc_im = {0.f, 0.f, 0.f, 0.f};
  int s = 0;
  while (s < 15 && start[s + 1] <= 4 * g0) ++s;
  int next = start[s + 1];
  auto flush = [&](int sbin) {
    const int x = (sbin & 3) * SB + i;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int cell = ((sbin >> 2) * SB + 4 * k + rr) * T + x;
      if constexpr (PK) {
        if (acc_re[rr] != 0.0f || acc_im[rr] != 0.0f)
          atomicAdd(&sre[cell], pack_re_im((double)acc_re[rr], (double)acc_im[rr], scale));
      } else {
        if (acc_re[rr] != 0.0f) atomicAdd(&sre[cell], TileAcc<float>::quant((double)acc_re[rr], scale));
        if (acc_im[rr] != 0.0f) atomicAdd(&sim[cell], TileAcc<float>::quant((double)acc_im[rr], scale));
      }
    }
  };
  // the records of the next batch are loaded while the current one is processed
  auto fetch = [&](int gb) {
    const int q = 4 * gb + sq;
    const int j = q < 4 * g1 ? order[q] : kPad;
    return j != kPad ? pk[j] : make_uint4(0u, 0u, 0u, 0u);
  };
  uint4 p = fetch(g0);
  for (int gb = g0; gb < g1 && finite; gb += kWinBatch / 4) {
    // -- weights of the batch's points
    const int q = 4 * gb + sq;
    if (q < 4 * g1) {
      const unsigned wd = sdim ? p.y : p.x;
      const float t = (float)((double)(wd & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
#pragma unroll
      for (int m = 0; m < NS / 2; ++m) myrow[m] = es_phi(t + (float)(sh * (NS / 2) + m), bl2e, kc);
      if (lane < kWinBatch) {
        const int lx = (int)(p.x >> 25), ly = (int)(p.y >> 25);
        const int ox = lx - min(lx / SB, 3) * SB, oy = ly - min(ly / SB, 3) * SB;  // offsets in the window
        // LDS float index of entry 16 - ns - o of the point's rows (a lane reads it + i)
        meta[sq] = make_int4(ry0 + kWinRow * sq - oy, rx0 + kWinRow * sq - ox, (int)p.z, (int)p.w);
      }
    }
    // the next batch's records load while this batch's groups run
    p = fetch(gb + kWinBatch / 4);
    wave_sync_win();
    // -- one MFMA pair per group of 4 points; all operands of the batch are read first
    // (entries past the wave's range hold stale but in-range read indices)
    constexpr int GB = kWinBatch / 4;
    const int ng = min(g1 - gb, GB);
    int4 mt[GB];
    float wy[GB], wx[GB];
#pragma unroll
    for (int u = 0; u < GB; ++u) mt[u] = meta[4 * u + k];
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      wy[u] = lf[mt[u].x + i];
      wx[u] = lf[mt[u].y + i];
    }
#pragma unroll
    for (int u = 0; u < GB; ++u) {
      if (u >= ng) break;
      const int g = gb + u;
      if (4 * g >= next) {
        flush(s);
        acc_re = win_f4{0.f, 0.f, 0.f, 0.f};
        acc_im = win_f4{0.f, 0.f, 0.f, 0.f};
        do {
          ++s;
        } while (s < 15 && start[s + 1] <= 4 * g);
        next = start[s + 1];
      }
      const float are = __int_as_float(mt[u].z) * wy[u], aim = __int_as_float(mt[u].w) * wy[u];
      acc_re = __builtin_amdgcn_mfma_f32_16x16x4f32(are, wx[u], acc_re, 0, 0, 0);
      acc_im = __builtin_amdgcn_mfma_f32_16x16x4f32(aim, wx[u], acc_im, 0, 0, 0);
    }
    wave_sync_win();
  }
  if (g1 > g0 && finite) flush(s);
  __syncthreads();

  // ---- 3. flush the tile's non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < TC; cell += kWinThreads) {
    long long ir, ii;
    if constexpr (PK) {
      unpack_re_im(sre[cell], ir, ii);
    } else {
      ir = (long long)sre[cell];
      ii = (long long)sim[cell];
    }
    if (finite && ir == 0 && ii == 0) continue;
    const float vr = finite ? (float)((double)ir * inv_scale) : NAN;
    const float vi = finite ? (float)((double)ii * inv_scale) : NAN;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// ---- fp16-split window spread
//
// The same windows, with the matrix products on v_mfma_f32_16x16x32_f16 (32 points per
// instruction, measured at 18 clocks on MI355X against 36 for v_mfma_f32_16x16x4f32 with
// 4 points; VALU and MFMA instructions of one SIMD do not run concurrently, so MFMA clocks
// add to the kernel time).  FP32 operands are split into two fp16 parts, v = hi + lo with
// hi = fp16(v) and lo = fp16(v - hi), and each window accumulates
//   A_hi B_hi + A_lo B_hi + A_hi B_lo          (three MFMAs per part, FP32 accumulation),
// which leaves out A_lo B_lo and the rounding of lo: at most 2^-22 |A| |B| per term.  Both
// operands are scaled to the top of fp16's range, so that lo parts stay normal numbers (the
// matrix cores flush fp16 subnormal inputs: with B = wx in [0, 1] the lo parts of weights
// below 2^-3 were lost, a 5e-6 relative error on the 2D radial workload):
// A = ldexp(c, 15 - e) wy with 2^e above the subproblem's max |re c|, |im c| (|A| < 2^15) and
// B = 2^14 wx (B < 2^15); lo is then normal down to |A|, B = 2^-3, i.e. relative to the
// largest operand 2^-18, and the terms lost below that are under 2^-29 of it.
//
// Subproblems hold up to 4096 points (8 per thread while ordering), so one block covers a
// bin of the 2D radial workload where the FP32 kernel needed four: the tile flush to the
// fine grid and the window flushes are paid four times less often.  Sub-bins are padded to
// multiples of 32 points (zero-strength entries).  Tiles are exact (two 64-bit words per
// cell, scale 2^(50 - e) with 2^e above 2 n max(|re c|, |im c|)).
constexpr int kWhThreads = 512;                  // 8 waves per subproblem
constexpr int kWhPpt = 8;                        // points per thread while ordering
constexpr int kWhMaxPts = kWhThreads * kWhPpt;   // 4096 points per subproblem
static_assert(kWhMaxPts == kWinMaxSub, "win2d_geometry admits subproblems of up to kWinMaxSub points");
constexpr int kWhK = 32;                         // points per group (K of the f16 MFMA)
constexpr int kWhRowF = kWhK * kWinRow + 16;     // floats of one row block (+ trailing zeros)
constexpr int kWhWaveF = 2 * kWhRowF + 4 * kWhK; // floats per wave: x rows, y rows, meta

typedef _Float16 wh_h8 __attribute__((ext_vector_type(8)));

__host__ __device__ inline size_t wh_order_bytes(int maxsub) {
  return ((size_t)(maxsub + 16 * (kWhK - 1)) * sizeof(uint16_t) + 15) / 16 * 16;
}

// LDS bytes of the fp16-split window spread (dynamic part) for subproblems of maxsub points
template <int NS>
inline size_t wh_spread_lds(int maxsub) {
  constexpr int T = WinGeom<NS>::T;
  return 2 * (size_t)T * T * sizeof(unsigned long long) + wh_order_bytes(maxsub) +
         (size_t)(kWhThreads / 64) * kWhWaveF * sizeof(float);
}

// v = hi + lo (fp16 parts).  lo must be taken from the very hi that the caller uses: the
// empty asm makes hi one opaque value, so that the compiler cannot rebuild it in another
// way for the subtraction (a hi from the exact product there and from the rounded one in the
// operand left hi + lo a whole fp16 ulp off now and then).
__device__ __forceinline__ void split_f16(const float (&v)[8], wh_h8& hi, wh_h8& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) hi[e] = (_Float16)v[e];
  asm("" : "+v"(hi));
#pragma unroll
  for (int e = 0; e < 8; ++e) lo[e] = (_Float16)(v[e] - (float)hi[e]);
}

// a b = hi + lo for the products of a and b (lo from the exact product, one rounding)
__device__ __forceinline__ void split_prod_f16(const float (&a)[8], const float (&b)[8], wh_h8& hi, wh_h8& lo) {
#pragma unroll
  for (int e = 0; e < 8; ++e) hi[e] = (_Float16)(a[e] * b[e]);
  asm("" : "+v"(hi));
#pragma unroll
  for (int e = 0; e < 8; ++e) lo[e] = (_Float16)__builtin_fmaf(a[e], b[e], -(float)hi[e]);
}

// POLY: kernel weights from the polynomial ES kernel (ns = 8, common.hpp), else es_phi_in
template <int NS, bool POLY>
__global__ void __launch_bounds__(kWhThreads) k_spread_win2dh(SIArgs<float> a, int maxsub,
                                                              const uint4* __restrict__ packed,
                                                              cplx<float>* __restrict__ fw) {
  using G = WinGeom<NS>;
  using Word = unsigned long long;
  constexpr int SB = G::SB, T = G::T, TC = T * T;
  constexpr int W = kWhThreads / 64;
  constexpr uint16_t kPad = 0xFFFF;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Word* sre = reinterpret_cast<Word*>(smem_raw);
  Word* sim = sre + TC;
  uint16_t* order = reinterpret_cast<uint16_t*>(sim + TC);
  __shared__ unsigned short wcnt[kWhPpt * W * 16];  // per (pass, wave, sub-bin): count, then first slot
  __shared__ int cnt[16], start[17];
  __shared__ unsigned smax;
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int4 sd = a.subprob[sub];
  const int n = sd.z;
  const uint4* __restrict__ pk = packed + (int64_t)tr * a.M + sd.y;
  float* const lf = reinterpret_cast<float*>(smem_raw);
  const int wof = (int)((2 * (size_t)TC * sizeof(Word) + wh_order_bytes(maxsub)) / 4) + wave * kWhWaveF;
  float* rowx = lf + wof;         // x weights: row q at 17 q, weights at 17 q + 16 - NS .. 17 q + 15
  float* rowy = rowx + kWhRowF;   // y weights, same layout
  int4* meta = reinterpret_cast<int4*>(rowy + kWhRowF);  // {y read index, x read index, re A, im A}
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  for (int e = tid; e < TC; e += kWhThreads) {
    sre[e] = 0;
    sim[e] = 0;
  }
  for (int e = lane; e < kWhRowF; e += 64)  // the zero parts of the weight rows stay zero
    if (e % kWinRow < 16 - NS || e % kWinRow == 16 || e >= kWhK * kWinRow) {
      rowx[e] = 0.0f;
      rowy[e] = 0.0f;
    }
  for (int e = tid; e < kWhPpt * W * 16; e += kWhThreads) wcnt[e] = 0;
  if (tid == 0) smax = 0u;
  __syncthreads();

  // ---- 1. order (as the FP32 kernel, 8 points per thread, sub-bins padded to 32)
  int sb[kWhPpt];
  unsigned mymax = 0;
#pragma unroll
  for (int r = 0; r < kWhPpt; ++r) {
    const int j = tid + r * kWhThreads;
    sb[r] = -1;
    if (j < n) {
      const uint4 p = pk[j];
      const int sx = min((int)(p.x >> 25) / SB, 3), sy = min((int)(p.y >> 25) / SB, 3);
      const float are = fabsf(__uint_as_float(p.z)), aim = fabsf(__uint_as_float(p.w));
      mymax = max(mymax, max(__float_as_uint(are), __float_as_uint(aim)));
      sb[r] = sy * 4 + sx;
    }
  }
  int rank[kWhPpt];
#pragma unroll
  for (int r = 0; r < kWhPpt; ++r) {
    const bool valid = sb[r] >= 0;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint64_t bal = __ballot(valid && ((sb[r] >> b) & 1));
      m &= ((sb[r] >> b) & 1) ? bal : ~bal;
    }
    rank[r] = popc_below(m);
    if (valid && rank[r] == 0) wcnt[(r * W + wave) * 16 + sb[r]] = (unsigned short)__popcll(m);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mymax = max(mymax, (unsigned)__shfl_xor((int)mymax, off, 64));
  if (lane == 0) atomicMax(&smax, mymax);
  __syncthreads();
  if (tid < 16) {
    int c = 0;
    for (int e = 0; e < kWhPpt * W; ++e) c += wcnt[e * 16 + tid];
    cnt[tid] = c;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int s = 0; s < 16; ++s) {
      start[s] = acc;
      acc += (cnt[s] + kWhK - 1) & ~(kWhK - 1);
    }
    start[16] = acc;
  }
  __syncthreads();
  if (tid < 16) {
    int run = start[tid];
    for (int e = 0; e < kWhPpt * W; ++e) {
      const int v = wcnt[e * 16 + tid];
      wcnt[e * 16 + tid] = (unsigned short)run;
      run += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kWhPpt; ++r)
    if (sb[r] >= 0) order[wcnt[(r * W + wave) * 16 + sb[r]] + rank[r]] = (uint16_t)(tid + r * kWhThreads);
  if (tid < 16 * (kWhK - 1)) {  // zero-strength padding entries
    const int s = tid / (kWhK - 1), q = tid % (kWhK - 1);
    const int slot = start[s] + cnt[s] + q;
    if (slot < start[s + 1]) order[slot] = kPad;
  }
  const float cm = __uint_as_float(smax);
  const double B = 2.0 * (double)n * (double)cm;
  const bool finite = isfinite(B) && isfinite(cm);
  double scale = 1.0, inv_scale = 1.0, wscale = 1.0;
  int ke = 0;  // strengths scale into fp16's range: A = ldexp(c, ke) wy, |A| < 2^15
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);  // B < 2^e
    scale = ldexp(1.0, 50 - e);
    inv_scale = ldexp(1.0, e - 50);
    int ec;
    (void)frexpf(cm, &ec);  // cm < 2^ec
    ke = 15 - ec;
    wscale = ldexp(scale, -ke - 14);  // window values are in units of 2^-(ke + 14)
  }
  __syncthreads();

  // ---- 2. accumulate: groups of 32 points of one sub-bin, contiguous ranges per wave
  const int NG = start[16] / kWhK;
  const int g0 = (int)(((int64_t)NG * wave) / W), g1 = (int)(((int64_t)NG * (wave + 1)) / W);
  const int i = lane & 15, kq = lane >> 4;
  const int sq = lane & (kWhK - 1), sdim = lane >> 5;  // weight staging: point sq, dimension sdim
  const float bl2e = a.es.beta_log2e, kc = a.es.c;
  float* const myrow = (sdim ? rowy : rowx) + kWinRow * sq + (16 - NS);
  const int ry0 = wof + kWhRowF + (16 - NS), rx0 = wof + (16 - NS);  // first weight of row 0 (y, x)
  // the matrix cores align each product to the largest addend and truncate it there, which
  // biases a long accumulation; so every group's three MFMAs start from zero (t) and the group
  // sums add into the window in FP32 on the VALU (round to nearest), one group later, when
  // the MFMAs have long finished
  const win_f4 zero4 = {0.f, 0.f, 0.f, 0.f};
  win_f4 acc_re = zero4, acc_im = zero4, t_re = zero4, t_im = zero4;
  int s = 0;
  while (s < 15 && start[s + 1] <= kWhK * g0) ++s;
  int next = start[s + 1];
  auto flush = [&](int sbin) {
    const int x = (sbin & 3) * SB + i;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int cell = ((sbin >> 2) * SB + 4 * kq + rr) * T + x;
      atomicAdd(&sre[cell], TileAcc<float>::quant((double)acc_re[rr], wscale));
      atomicAdd(&sim[cell], TileAcc<float>::quant((double)acc_im[rr], wscale));
    }
  };
  auto fetch = [&](int g) {
    const int j = g < g1 ? order[kWhK * g + sq] : kPad;
    return j != kPad ? pk[j] : make_uint4(0u, 0u, 0u, 0u);
  };
  uint4 p = fetch(g0);
  for (int g = g0; g < g1 && finite; ++g) {
    acc_re += t_re;
    acc_im += t_im;
    if (kWhK * g >= next) {  // the group opens the next non-empty sub-bin
      flush(s);
      acc_re = zero4;
      acc_im = zero4;
      do {
        ++s;
      } while (s < 15 && start[s + 1] <= kWhK * g);
      next = start[s + 1];
    }
    // -- kernel weights of the group's points (lanes 0-31 x, 32-63 y)
    {
      const unsigned wd = sdim ? p.y : p.x;
      const float wsc = sdim ? 1.0f : 16384.0f;  // B = 2^14 wx
      if constexpr (POLY) {
        // padding entries (word 0, v = -3, zero strength) take v = -1: finite weights
        float w[NS];
        es_poly_weights<NS>(fmaxf(es_poly_v(wd), -1.0f), a.poly, w);
#pragma unroll
        for (int m = 0; m < NS; ++m) myrow[m] = w[m] * wsc;
      } else {
        const float t = (float)((double)(wd & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
        // t + m lies in [-NS/2, NS/2] for the points (t in [-NS/2, 1 - NS/2)); padding entries
        // (t = -NS/2 - 1, zero strength) get finite weights
#pragma unroll
        for (int m = 0; m < NS; ++m) myrow[m] = es_phi_in(t + (float)m, bl2e, kc) * wsc;
      }
      if (sdim == 0) {
        const int lx = (int)(p.x >> 25), ly = (int)(p.y >> 25);
        const int ox = lx - min(lx / SB, 3) * SB, oy = ly - min(ly / SB, 3) * SB;  // offsets in the window
        meta[sq] = make_int4(ry0 + kWinRow * sq - oy, rx0 + kWinRow * sq - ox,
                             __float_as_int(ldexpf(__uint_as_float(p.z), ke)),
                             __float_as_int(ldexpf(__uint_as_float(p.w), ke)));
      }
    }
    p = fetch(g + 1);
    wave_sync_win();
    // -- operands: lane (i, kq) holds A[i][8 kq + e] and B[8 kq + e][i]
    float cre[8], cim[8], wy[8], bx[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int4 mt = meta[8 * kq + e];
      wy[e] = lf[mt.x + i];
      bx[e] = lf[mt.y + i];
      cre[e] = __int_as_float(mt.z);
      cim[e] = __int_as_float(mt.w);
    }
    wave_sync_win();
    wh_h8 bh, bl, ah, al;
    split_f16(bx, bh, bl);
    split_prod_f16(cre, wy, ah, al);
    t_re = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, zero4, 0, 0, 0);
    t_re = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, t_re, 0, 0, 0);
    t_re = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, t_re, 0, 0, 0);
    split_prod_f16(cim, wy, ah, al);
    t_im = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, zero4, 0, 0, 0);
    t_im = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, t_im, 0, 0, 0);
    t_im = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, t_im, 0, 0, 0);
  }
  acc_re += t_re;
  acc_im += t_im;
  if (g1 > g0 && finite) flush(s);
  __syncthreads();

  // ---- 3. flush the tile's non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  const int tile3[3] = {T, T, 1};
  for (int cell = tid; cell < TC; cell += kWhThreads) {
    const long long ir = (long long)sre[cell], ii = (long long)sim[cell];
    if (finite && ir == 0 && ii == 0) continue;
    const float vr = finite ? (float)((double)ir * inv_scale) : NAN;
    const float vi = finite ? (float)((double)ii * inv_scale) : NAN;
    const int64_t gi = tile_cell_to_global<2>(cell, tile3, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// ---- fp16-split window interpolation (2D FP32, the transpose of the window spread)
//
// The values of the points of one sub-bin read one 16 x 16 window F of the tile:
//   v_j = sum_y wy_j(y) sum_x F[y][x] wx_j(x)                (nufftax/core/spread.py:596-658)
// so for a group of 16 points G = F Bx with Bx[x][j] = wx_j(x) is one 16 x 16 x 16 matrix
// product (v_mfma_f32_16x16x16_f16, three per part with fp16-split operands as in the
// spread), and v_j = sum_y wy_j(y) G[y][j] is a 16-term dot product: each lane holds four y
// of one point, two cross-lane adds finish it.  The tile is loaded once per subproblem; a
// window is split into fp16 parts once per sub-bin a wave visits.  Scaling: A = ldexp(F,
// 15 - e) with 2^e above the tile's max |re|, |im|, B = 2^14 wx.  A non-finite tile value
// makes the values of the points of its window non-finite (the product meets it with a zero
// weight).  Points come from the compact records (sorted order, perm gives the output index).
constexpr int kWiThreads = 512;                  // 8 waves per subproblem
constexpr int kWiPpt = 8;                        // points per thread while ordering (<= 4096)
constexpr int kWiK = 16;                         // points per group (K of the f16 MFMA's output columns)
constexpr int kWiRowF = kWiK * kWinRow + 16;     // floats of one row block (+ trailing zeros)
constexpr int kWiWaveF = 2 * kWiRowF + 4 * kWiK; // floats per wave: x rows, y rows, meta
static_assert(kWiThreads * kWiPpt == kWinMaxSub, "the interpolation orders subproblems of up to kWinMaxSub points");

typedef _Float16 wi_h4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline size_t wi_order_bytes(int maxsub) {
  return ((size_t)(maxsub + 16 * (kWiK - 1)) * sizeof(uint16_t) + 15) / 16 * 16;
}

// LDS bytes of the window interpolation (dynamic part) for subproblems of maxsub points
template <int NS>
inline size_t wi_interp_lds(int maxsub) {
  constexpr int T = WinGeom<NS>::T;
  return (size_t)T * T * sizeof(cplx<float>) + wi_order_bytes(maxsub) + (size_t)(kWiThreads / 64) * kWiWaveF * sizeof(float);
}

__device__ __forceinline__ void split4_f16(const float (&v)[4], wi_h4& hi, wi_h4& lo) {
#pragma unroll
  for (int e = 0; e < 4; ++e) hi[e] = (_Float16)v[e];
  asm("" : "+v"(hi));  // one hi for both uses (see split_f16)
#pragma unroll
  for (int e = 0; e < 4; ++e) lo[e] = (_Float16)(v[e] - (float)hi[e]);
}

template <int NS>
__global__ void __launch_bounds__(kWiThreads) k_interp_win2dh(SIArgs<float> a, int maxsub,
                                                              const cplx<float>* __restrict__ fw,
                                                              cplx<float>* __restrict__ out) {
  using G = WinGeom<NS>;
  constexpr int SB = G::SB, T = G::T, TC = T * T;
  constexpr int W = kWiThreads / 64;
  constexpr uint16_t kPad = 0xFFFF;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<float>* tile = reinterpret_cast<cplx<float>*>(smem_raw);
  uint16_t* order = reinterpret_cast<uint16_t*>(tile + TC);
  __shared__ unsigned short wcnt[kWiPpt * W * 16];
  __shared__ int cnt[16], start[17];
  __shared__ unsigned smax;
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int4 sd = a.subprob[sub];
  const int n = sd.z;
  const uint2* __restrict__ rec = a.crec + sd.y;
  float* const lf = reinterpret_cast<float*>(smem_raw);
  const int wof = (int)(((size_t)TC * sizeof(cplx<float>) + wi_order_bytes(maxsub)) / 4) + wave * kWiWaveF;
  float* rowx = lf + wof;
  float* rowy = rowx + kWiRowF;
  int4* meta = reinterpret_cast<int4*>(rowy + kWiRowF);  // {y read index, x read index, output index, -}
  int org[3];
  tile_origin<2>(sd, a.bs, a.nb, NS / 2, org);
  const int tile3[3] = {T, T, 1};
  const cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  if (tid == 0) smax = 0u;
  for (int e = tid; e < kWiPpt * W * 16; e += kWiThreads) wcnt[e] = 0;
  for (int e = lane; e < kWiRowF; e += 64)
    if (e % kWinRow < 16 - NS || e % kWinRow == 16 || e >= kWiK * kWinRow) {
      rowx[e] = 0.0f;
      rowy[e] = 0.0f;
    }
  __syncthreads();
  // ---- 1. tile and its max |re|, |im|
  unsigned mymax = 0;
  for (int cell = tid; cell < TC; cell += kWiThreads) {
    const cplx<float> v = fT[tile_cell_to_global<2>(cell, tile3, org, a.nf)];
    tile[cell] = v;
    mymax = max(mymax, max(__float_as_uint(fabsf(v.re)), __float_as_uint(fabsf(v.im))));
  }
  // ---- order the points by sub-bin (as in the window spread, groups of 16)
  int sb[kWiPpt];
  const int npass = (n + kWiThreads - 1) / kWiThreads;  // passes that hold points (uniform)
#pragma unroll
  for (int r = 0; r < kWiPpt; ++r) {
    const int j = tid + r * kWiThreads;
    sb[r] = -1;
    if (j < n) {
      const uint2 p = rec[j];
      sb[r] = min((int)(p.y >> 25) / SB, 3) * 4 + min((int)(p.x >> 25) / SB, 3);
    }
  }
  int rank[kWiPpt];
#pragma unroll
  for (int r = 0; r < kWiPpt; ++r) {
    if (r >= npass) break;
    const bool valid = sb[r] >= 0;
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const uint64_t bal = __ballot(valid && ((sb[r] >> b) & 1));
      m &= ((sb[r] >> b) & 1) ? bal : ~bal;
    }
    rank[r] = popc_below(m);
    if (valid && rank[r] == 0) wcnt[(r * W + wave) * 16 + sb[r]] = (unsigned short)__popcll(m);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) mymax = max(mymax, (unsigned)__shfl_xor((int)mymax, off, 64));
  if (lane == 0) atomicMax(&smax, mymax);
  __syncthreads();
  if (tid < 16) {
    int c = 0;
    for (int e = 0; e < kWiPpt * W; ++e) c += wcnt[e * 16 + tid];
    cnt[tid] = c;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int s = 0; s < 16; ++s) {
      start[s] = acc;
      acc += (cnt[s] + kWiK - 1) & ~(kWiK - 1);
    }
    start[16] = acc;
  }
  __syncthreads();
  if (tid < 16) {
    int run = start[tid];
    for (int e = 0; e < kWiPpt * W; ++e) {
      const int v = wcnt[e * 16 + tid];
      wcnt[e * 16 + tid] = (unsigned short)run;
      run += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kWiPpt; ++r)
    if (r < npass && sb[r] >= 0) order[wcnt[(r * W + wave) * 16 + sb[r]] + rank[r]] = (uint16_t)(tid + r * kWiThreads);
  if (tid < 16 * (kWiK - 1)) {
    const int s = tid / (kWiK - 1), q = tid % (kWiK - 1);
    const int slot = start[s] + cnt[s] + q;
    if (slot < start[s + 1]) order[slot] = kPad;
  }
  const float fm = __uint_as_float(smax);
  int ke = 0;  // A = ldexp(F, ke), |A| < 2^15
  if (isfinite(fm) && fm > 0.0f) {
    int ef;
    (void)frexpf(fm, &ef);
    ke = 15 - ef;
  }
  const int kout = -ke - 14;  // values come out in units of 2^(ke + 14)
  __syncthreads();

  // ---- 2. groups of 16 points of one sub-bin, contiguous ranges per wave
  const int NG = start[16] / kWiK;
  const int g0 = (int)(((int64_t)NG * wave) / W), g1 = (int)(((int64_t)NG * (wave + 1)) / W);
  const int i = lane & 15, kq = lane >> 4;
  // weight staging: lane -> (point sq, dimension sdim, weights sh NS/2 .. sh NS/2 + NS/2 - 1)
  const int sq = lane & (kWiK - 1), sdim = (lane >> 4) & 1, sh = lane >> 5;
  const float bl2e = a.es.beta_log2e, kc = a.es.c;
  float* const myrow = (sdim ? rowy : rowx) + kWinRow * sq + (16 - NS) + sh * (NS / 2);
  const int ry0 = wof + kWiRowF + (16 - NS), rx0 = wof + (16 - NS);
  int s = -1;  // sub-bin whose window is split in registers
  int next = 0;
  wi_h4 fre_h, fre_l, fim_h, fim_l;
  // the next group's record and output index load while the current group runs
  auto fetch = [&](int g, uint2& pr, int& outi) {
    const int j = g < g1 ? (int)order[kWiK * g + sq] : (int)kPad;
    pr = j != kPad ? rec[j] : make_uint2(0u, 0u);
    outi = j != kPad ? a.perm[sd.y + j] : -1;
  };
  uint2 p;
  int outi;
  fetch(g0, p, outi);
  for (int g = g0; g < g1; ++g) {
    if (kWiK * g >= next) {  // a new sub-bin: its window, split into fp16 parts
      do {
        ++s;
      } while (s < 15 && start[s + 1] <= kWiK * g);
      next = start[s + 1];
      // A[i][k] = F[y = i][x = 4 kq + e]: four consecutive cells of tile row sy SB + i
      const cplx<float>* row = tile + ((s >> 2) * SB + i) * T + (s & 3) * SB + 4 * kq;
      float vr[4], vi[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const cplx<float> v = row[e];
        vr[e] = ldexpf(v.re, ke);
        vi[e] = ldexpf(v.im, ke);
      }
      split4_f16(vr, fre_h, fre_l);
      split4_f16(vi, fim_h, fim_l);
    }
    {
      const unsigned wd = sdim ? p.y : p.x;
      const float t = (float)((double)(wd & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
      const float wsc = sdim ? 1.0f : 16384.0f;  // B = 2^14 wx
#pragma unroll
      for (int m = 0; m < NS / 2; ++m) myrow[m] = es_phi_in(t + (float)(sh * (NS / 2) + m), bl2e, kc) * wsc;
      if (lane < kWiK) {
        const int lx = (int)(p.x >> 25), ly = (int)(p.y >> 25);
        const int ox = lx - min(lx / SB, 3) * SB, oy = ly - min(ly / SB, 3) * SB;
        meta[sq] = make_int4(ry0 + kWinRow * sq - oy, rx0 + kWinRow * sq - ox, outi, 0);
      }
    }
    fetch(g + 1, p, outi);
    wave_sync_win();
    // B[k = x][j] in lane j + 16 (x / 4): point i, x = 4 kq + e; wy for y = 4 kq + e likewise
    const int4 mt = meta[i];
    float bx[4], wy[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      bx[e] = lf[mt.y + 4 * kq + e];
      wy[e] = lf[mt.x + 4 * kq + e];
    }
    wave_sync_win();
    wi_h4 bh, bl;
    split4_f16(bx, bh, bl);
    const win_f4 z = {0.f, 0.f, 0.f, 0.f};
    win_f4 gre = __builtin_amdgcn_mfma_f32_16x16x16f16(fre_l, bh, z, 0, 0, 0);
    gre = __builtin_amdgcn_mfma_f32_16x16x16f16(fre_h, bl, gre, 0, 0, 0);
    gre = __builtin_amdgcn_mfma_f32_16x16x16f16(fre_h, bh, gre, 0, 0, 0);
    win_f4 gim = __builtin_amdgcn_mfma_f32_16x16x16f16(fim_l, bh, z, 0, 0, 0);
    gim = __builtin_amdgcn_mfma_f32_16x16x16f16(fim_h, bl, gim, 0, 0, 0);
    gim = __builtin_amdgcn_mfma_f32_16x16x16f16(fim_h, bh, gim, 0, 0, 0);
    // G[y = 4 kq + e][point i] in element e: dot with wy, then over the four lane groups
    float sr = 0.0f, si = 0.0f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      sr = fmaf(gre[e], wy[e], sr);
      si = fmaf(gim[e], wy[e], si);
    }
    sr += __shfl_xor(sr, 16, 64);
    si += __shfl_xor(si, 16, 64);
    sr += __shfl_xor(sr, 32, 64);
    si += __shfl_xor(si, 32, 64);
    if (lane < kWiK && mt.z >= 0 && mt.z < a.M)
      out[(int64_t)tr * a.M + mt.z] = cplx<float>{ldexpf(sr, kout), ldexpf(si, kout)};
  }
}


// NUFFTX_WIN_F32=1: the FP32 window kernel (v_mfma_f32_16x16x4f32, subproblems of at most
// 1024 points) instead of the fp16-split one
inline bool win_f32_kernel() {
  const char* e = std::getenv("NUFFTX_WIN_F32");
  return e != nullptr && e[0] == '1';
}

// NUFFTX_WIN_EXACT=1: two 64-bit words per tile cell (exact tile sums) instead of packed tiles
inline bool win_exact_tiles() {
  const char* e = std::getenv("NUFFTX_WIN_EXACT");
  return e != nullptr && e[0] == '1';
}


}  // namespace nufftx
