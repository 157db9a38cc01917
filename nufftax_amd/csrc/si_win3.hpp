// 3D FP32 spreading on the matrix cores: the window spread of si_win.hpp, one 16 x 16 (x, y)
// window per z slice.
//
// A point's contribution c wz(z) wy(y) wx(x) (nufftax/core/spread.py:465-489) is, for each z
// slice of its support, an outer product; the points of one sub-bin of 8 x 8 x 4 support
// starts land in a window of 16 x 16 cells in (x, y) and WZ = 4 + ns slices in z, so a group
// of 32 points is one v_mfma_f32_16x16x32_f16 product per slice and part:
//   W_z[y][x] += sum_k (c_k wy_k(y)) (wz_k(z) wx_k(x))       A = c wy (split once per group),
//                                                             B_z = wz(z) wx (split per slice)
// with fp16-split operands (A_hi B_hi + A_lo B_hi + A_hi B_lo, at most 2^-22 |A| |B| per term
// off) accumulating in FP32 registers, WZ windows per part.  Sub-bins are 8 support starts
// wide in x and y whatever ns (the window reaches 8 + ns <= 16 cells; cells past the tile,
// which no support reaches, are never flushed) and 4 in z.  Bins must be whole numbers of
// sub-bins (at most 16 per bin); the 3D sort's bins of 16 x 8 x 8 or 16 x 16 x 8 cells are.
//
// Block = one subproblem (<= 2048 points of one bin), 8 waves:
//   1. order: as in the 2D kernel (four points per thread), sub-bins padded to multiples of
//      32 with zero-strength entries; the block also takes max(|re c|, |im c|) and
//      S = sum_j max(|re c_j|, |im c_j|);
//   2. accumulate: contiguous ranges of groups per wave; per group, the lanes evaluate the
//      kernel weights of the 32 points into LDS rows (x and y: 17 floats a row, 16 - ns zeros
//      first; z: 13 floats a row, 4 zeros first), then each lane reads its operands for 8
//      points, splits A once and B once per slice, and issues 6 MFMAs per slice.  When a
//      wave's walk leaves a sub-bin, its windows go into the block's tile, in fixed point:
//      exact tiles (PK false) hold two 64-bit words per cell at scale 2^(50 - e) with 2^e
//      above 2 n max|c|, which bounds every partial sum, so the tile's sums do not depend on
//      the order of the adds; packed tiles (PK true) hold one word per cell (real part high,
//      imaginary low, scale 2^(30 - e) with 2^e above S, which bounds every window value and
//      every final tile value; the words add modulo 2^64 and only final sums must fit 31 bits);
//   3. flush: the tile's non-zero cells go to the fine grid with FP32 global atomics.
// The per-wave LDS blocks (weight rows, meta, strengths) are whole multiples of 16 bytes: an
// odd-sized block misaligns the int4 meta reads of every other wave (1.70 -> 1.93 ms on 8M
// clustered points).
#pragma once

#include "si_win.hpp"

namespace nufftx {

constexpr int kW3Threads = 512;                    // 8 waves per subproblem
constexpr int kW3Ppt = 4;                          // points per thread while ordering
constexpr int kW3K = 32;                           // points per group
constexpr int kW3SB = 8;                           // sub-bin side in x and y (support starts)
constexpr int kW3SBZ = 4;                          // sub-bin side in z
constexpr int kW3RowZ = 13;                        // z row stride (floats): 4 zeros, ns <= 8 weights, 1 zero
// row blocks: the last row's reads reach 16 - ns + 15 past its start (x, y) and 4 + 3 + ns (z);
// every block is a whole number of 16 bytes, so that each wave's int4 meta and float2
// strengths stay aligned
constexpr int kW3RowF = ((kW3K - 1) * kWinRow + 32 + 3) & ~3;   // x / y row block (floats)
constexpr int kW3RowZF = ((kW3K - 1) * kW3RowZ + 16 + 3) & ~3;  // z row block (floats)
constexpr int kW3WaveF = 2 * kW3RowF + kW3RowZF + 4 * kW3K + 2 * kW3K;  // + meta int4 + strength float2
static_assert(kW3RowF % 4 == 0 && kW3RowZF % 4 == 0 && kW3WaveF % 4 == 0, "16-byte aligned per-wave blocks");
static_assert(kW3Threads * kW3Ppt == kWin3MaxSub, "win3d_geometry admits subproblems of up to kWin3MaxSub points");

__host__ __device__ inline size_t w3_order_bytes(int maxsub) {
  return ((size_t)(maxsub + 16 * (kW3K - 1)) * sizeof(uint16_t) + 15) / 16 * 16;
}

// LDS bytes of the 3D window spread (dynamic part); PK: one 64-bit word per tile cell, else two
template <bool PK>
inline size_t w3_spread_lds(int64_t tile_cells, int maxsub) {
  return (PK ? 1 : 2) * (size_t)tile_cells * sizeof(unsigned long long) + w3_order_bytes(maxsub) +
         (size_t)(kW3Threads / 64) * kW3WaveF * sizeof(float);
}

template <int NS, bool PK>
__global__ void __launch_bounds__(kW3Threads) k_spread_win3dh(SIArgs<float> a, int maxsub,
                                                              const uint4* __restrict__ words,
                                                              const float2* __restrict__ cs,
                                                              cplx<float>* __restrict__ fw) {
  using Word = unsigned long long;
  constexpr int WZ = kW3SBZ + NS;  // z slices of a window
  constexpr int W = kW3Threads / 64;
  constexpr uint16_t kPad = 0xFFFF;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int T0 = a.tile[0], T1 = a.tile[1], T2 = a.tile[2];
  const int TC = T0 * T1 * T2;
  // sub-bins per bin: nsx nsy nsz <= 16 (win3d_geometry)
  const int nsx = a.bs[0] / kW3SB, nsy = a.bs[1] / kW3SB, nsz = a.bs[2] / kW3SBZ;
  Word* tile = reinterpret_cast<Word*>(smem_raw);  // packed words, or real parts
  Word* tim = tile + TC;                            // imaginary parts (exact tiles)
  uint16_t* order = reinterpret_cast<uint16_t*>(tile + (PK ? 1 : 2) * TC);
  __shared__ unsigned short wcnt[kW3Ppt * W * 16];
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
  const uint4* __restrict__ pw = words + sd.y;
  const float2* __restrict__ pc = cs + (int64_t)tr * a.M + sd.y;
  float* const lf = reinterpret_cast<float*>(smem_raw);
  const int wof = (int)(((PK ? 1 : 2) * (size_t)TC * sizeof(Word) + w3_order_bytes(maxsub)) / 4) + wave * kW3WaveF;
  float* rowx = lf + wof;
  float* rowy = rowx + kW3RowF;
  float* rowz = rowy + kW3RowF;
  int4* meta = reinterpret_cast<int4*>(rowz + kW3RowZF);     // {y read index, x read index, z read index, -}
  float2* cval = reinterpret_cast<float2*>(meta + kW3K);      // scaled strengths
  int org[3];
  tile_origin<3>(sd, a.bs, a.nb, NS / 2, org);
  for (int e = tid; e < (PK ? 1 : 2) * TC; e += kW3Threads) tile[e] = 0;
  for (int e = lane; e < kW3RowF; e += 64)
    if (e % kWinRow < 16 - NS || e % kWinRow == 16 || e >= kW3K * kWinRow) {
      rowx[e] = 0.0f;
      rowy[e] = 0.0f;
    }
  for (int e = lane; e < kW3RowZF; e += 64)
    if (e % kW3RowZ < kW3SBZ || e % kW3RowZ >= kW3SBZ + NS || e >= kW3K * kW3RowZ) rowz[e] = 0.0f;
  for (int e = tid; e < kW3Ppt * W * 16; e += kW3Threads) wcnt[e] = 0;
  if (tid == 0) smax = 0u;
  __syncthreads();

  // ---- 1. order
  auto subbin = [&](const uint4& p) {
    const int sx = min((int)(p.x >> 25) / kW3SB, nsx - 1), sy = min((int)(p.y >> 25) / kW3SB, nsy - 1);
    const int sz = min((int)(p.z >> 25) / kW3SBZ, nsz - 1);
    return (sz * nsy + sy) * nsx + sx;
  };
  int sb[kW3Ppt];
  unsigned mymax = 0;
  float mysum = 0.0f;
  const int npass = (n + kW3Threads - 1) / kW3Threads;
#pragma unroll
  for (int r = 0; r < kW3Ppt; ++r) {
    const int j = tid + r * kW3Threads;
    sb[r] = -1;
    if (j < n) {
      sb[r] = subbin(pw[j]);
      const float2 v = pc[j];
      const float are = fabsf(v.x), aim = fabsf(v.y);
      mymax = max(mymax, max(__float_as_uint(are), __float_as_uint(aim)));
      mysum += fmaxf(are, aim) + 0.0f * (are + aim);  // NaN in either part propagates
    }
  }
  int rank[kW3Ppt];
#pragma unroll
  for (int r = 0; r < kW3Ppt; ++r) {
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
  for (int off = 32; off > 0; off >>= 1) {
    mymax = max(mymax, (unsigned)__shfl_xor((int)mymax, off, 64));
    mysum += __shfl_xor(mysum, off, 64);
  }
  if (lane == 0) {
    atomicMax(&smax, mymax);
    ssum[wave] = mysum;
  }
  __syncthreads();
  if (tid < 16) {
    int c = 0;
    for (int e = 0; e < kW3Ppt * W; ++e) c += wcnt[e * 16 + tid];
    cnt[tid] = c;
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int s = 0; s < 16; ++s) {
      start[s] = acc;
      acc += (cnt[s] + kW3K - 1) & ~(kW3K - 1);
    }
    start[16] = acc;
  }
  __syncthreads();
  if (tid < 16) {
    int run = start[tid];
    for (int e = 0; e < kW3Ppt * W; ++e) {
      const int v = wcnt[e * 16 + tid];
      wcnt[e * 16 + tid] = (unsigned short)run;
      run += v;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kW3Ppt; ++r)
    if (r < npass && sb[r] >= 0) order[wcnt[(r * W + wave) * 16 + sb[r]] + rank[r]] = (uint16_t)(tid + r * kW3Threads);
  if (tid < 16 * (kW3K - 1)) {
    const int s = tid / (kW3K - 1), q = tid % (kW3K - 1);
    const int slot = start[s] + cnt[s] + q;
    if (slot < start[s + 1]) order[slot] = kPad;
  }
  const float cm = __uint_as_float(smax);
  float S = 0.0f;
#pragma unroll
  for (int w = 0; w < W; ++w) S += ssum[w];
  // packed: S bounds every window value and tile sum (31 bits each); exact: 2 n max|c| bounds
  // every partial sum (50 bits)
  const double B = PK ? (double)S * (1.0 + 1.0 / 1024.0) + (double)cm : 2.0 * (double)n * (double)cm;
  const bool finite = isfinite(B) && isfinite(cm) && isfinite(S);
  constexpr int kBits = PK ? 30 : 50;
  double scale = 1.0, inv_scale = 1.0, wscale = 1.0;
  int ke = 0;  // A = ldexp(c, ke) wy, |A| < 2^15; B = 2^14 wx wz
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);
    scale = ldexp(1.0, kBits - e);
    inv_scale = ldexp(1.0, e - kBits);
    int ec;
    (void)frexpf(cm, &ec);
    ke = 15 - ec;
    wscale = ldexp(scale, -ke - 14);
  }
  __syncthreads();

  // ---- 2. accumulate
  const int NG = start[16] / kW3K;
  const int g0 = (int)(((int64_t)NG * wave) / W), g1 = (int)(((int64_t)NG * (wave + 1)) / W);
  const int i = lane & 15, kq = lane >> 4;
  const int sq = lane & (kW3K - 1), shi = lane >> 5;  // weight staging: point sq, x (0) or y (1); z half
  const float bl2e = a.es.beta_log2e, kc = a.es.c;
  const int ry0 = wof + kW3RowF + (16 - NS), rx0 = wof + (16 - NS), rz0 = wof + 2 * kW3RowF + kW3SBZ;
  win_f4 acc_re[WZ], acc_im[WZ];
#pragma unroll
  for (int z = 0; z < WZ; ++z) {
    acc_re[z] = win_f4{0.f, 0.f, 0.f, 0.f};
    acc_im[z] = win_f4{0.f, 0.f, 0.f, 0.f};
  }
  int s = 0;
  while (s < 15 && start[s + 1] <= kW3K * g0) ++s;
  int next = start[s + 1];
  auto flush = [&](int sbin) {
    const int sx = sbin % nsx, sy = (sbin / nsx) % nsy, sz = sbin / (nsx * nsy);
    const int x = sx * kW3SB + i;
#pragma unroll
    for (int z = 0; z < WZ; ++z) {
      const int zz = sz * kW3SBZ + z;
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int y = sy * kW3SB + 4 * kq + rr;
        // window cells past the tile hold nothing (no support reaches them)
        if (x >= T0 || y >= T1 || zz >= T2) continue;
        const int cell = (zz * T1 + y) * T0 + x;
        const float vr = acc_re[z][rr], vi = acc_im[z][rr];
        if constexpr (PK) {
          atomicAdd(&tile[cell], pack_re_im((double)vr, (double)vi, wscale));
        } else {
          atomicAdd(&tile[cell], TileAcc<float>::quant((double)vr, wscale));
          atomicAdd(&tim[cell], TileAcc<float>::quant((double)vi, wscale));
        }
      }
    }
  };
  auto fetch = [&](int g, uint4& pr, float2& cr) {
    const int j = g < g1 ? (int)order[kW3K * g + sq] : (int)kPad;
    pr = j != kPad ? pw[j] : make_uint4(0u, 0u, 0u, 0u);
    cr = j != kPad ? pc[j] : make_float2(0.0f, 0.0f);
  };
  uint4 p;
  float2 cv;
  fetch(g0, p, cv);
  for (int g = g0; g < g1 && finite; ++g) {
    if (kW3K * g >= next) {
      flush(s);
#pragma unroll
      for (int z = 0; z < WZ; ++z) {
        acc_re[z] = win_f4{0.f, 0.f, 0.f, 0.f};
        acc_im[z] = win_f4{0.f, 0.f, 0.f, 0.f};
      }
      do {
        ++s;
      } while (s < 15 && start[s + 1] <= kW3K * g);
      next = start[s + 1];
    }
    // -- kernel weights: lanes 0-31 x, 32-63 y (all ns); then z, ns/2 per lane
    {
      const unsigned wd = shi ? p.y : p.x;
      const float t = (float)((double)(wd & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
      const float wsc = shi ? 1.0f : 16384.0f;  // B carries 2^14
      float* const myrow = (shi ? rowy : rowx) + kWinRow * sq + (16 - NS);
#pragma unroll
      for (int m = 0; m < NS; ++m) myrow[m] = es_phi_in(t + (float)m, bl2e, kc) * wsc;
      const float tz = (float)((double)(p.z & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
      float* const zrow = rowz + kW3RowZ * sq + kW3SBZ + shi * (NS / 2);
#pragma unroll
      for (int m = 0; m < NS / 2; ++m) zrow[m] = es_phi_in(tz + (float)(shi * (NS / 2) + m), bl2e, kc);
      if (shi == 0) {
        const int lx = (int)(p.x >> 25), ly = (int)(p.y >> 25), lz = (int)(p.z >> 25);
        const int ox = lx - min(lx / kW3SB, nsx - 1) * kW3SB, oy = ly - min(ly / kW3SB, nsy - 1) * kW3SB;
        const int oz = lz - min(lz / kW3SBZ, nsz - 1) * kW3SBZ;
        meta[sq] = make_int4(ry0 + kWinRow * sq - oy, rx0 + kWinRow * sq - ox, rz0 + kW3RowZ * sq - oz, 0);
        cval[sq] = make_float2(ldexpf(cv.x, ke), ldexpf(cv.y, ke));
      }
    }
    fetch(g + 1, p, cv);
    wave_sync_win();
    // -- operands: lane (i, kq) holds A[i][8 kq + e] and B_z[8 kq + e][i]
    float cre[8], cim[8], wy[8], wx[8];
    int zb[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int4 mt = meta[8 * kq + e];
      const float2 c2 = cval[8 * kq + e];
      wy[e] = lf[mt.x + i];
      wx[e] = lf[mt.y + i];
      zb[e] = mt.z;
      cre[e] = c2.x;
      cim[e] = c2.y;
    }
    wh_h8 arh, arl, aih, ail;
    split_prod_f16(cre, wy, arh, arl);
    split_prod_f16(cim, wy, aih, ail);
#pragma unroll
    for (int z = 0; z < WZ; ++z) {
      float wz[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) wz[e] = lf[zb[e] + z];
      wh_h8 bh, bl;
      split_prod_f16(wx, wz, bh, bl);
      acc_re[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arl, bh, acc_re[z], 0, 0, 0);
      acc_re[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bl, acc_re[z], 0, 0, 0);
      acc_re[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(arh, bh, acc_re[z], 0, 0, 0);
      acc_im[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ail, bh, acc_im[z], 0, 0, 0);
      acc_im[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bl, acc_im[z], 0, 0, 0);
      acc_im[z] = __builtin_amdgcn_mfma_f32_16x16x32_f16(aih, bh, acc_im[z], 0, 0, 0);
    }
    wave_sync_win();
  }
  if (g1 > g0 && finite) flush(s);
  __syncthreads();

  // ---- 3. flush the tile's non-zero cells to the fine grid
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int cell = tid; cell < TC; cell += kW3Threads) {
    long long ir, ii;
    if constexpr (PK) {
      unpack_re_im(tile[cell], ir, ii);
    } else {
      ir = (long long)tile[cell];
      ii = (long long)tim[cell];
    }
    if (finite && ir == 0 && ii == 0) continue;
    const float vr = finite ? (float)((double)ir * inv_scale) : NAN;
    const float vi = finite ? (float)((double)ii * inv_scale) : NAN;
    const int64_t gi = tile_cell_to_global<3>(cell, a.tile, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

}  // namespace nufftx
