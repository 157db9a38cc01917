// Spreading (type 1) and interpolation (type 2) kernels on bin-sorted points.
//
// Math: fw[t, i] += c[t, j] * prod_d phi(i_d - xs_jd)   (nufftax/core/spread.py:435-494)
//       c[t, j]   = sum_i fw[t, i] * prod_d phi(...)     (spread.py:497-540)
// with i_d running over i0_d .. i0_d + ns - 1 (periodic), i0 = ceil(xs - ns/2).
//
// Tiled path: one workgroup per subproblem (<= maxsub points of one bin).  The
// bin plus an ns-wide halo is a tile in LDS; spreading accumulates with LDS
// atomics and flushes non-zero cells to HBM with global atomics, interpolation
// stages the tile once and gathers from LDS.  Weights are evaluated per point
// in registers (FP64 fold, T-precision ES kernel) and never stored.
// Global path (tile larger than the LDS budget): same per-point work directly
// on the fine grid in HBM, points visited in sorted order.
#pragma once
#include <cstdlib>

#include <type_traits>

#include "internal.hpp"

namespace nufftx {

template <typename T>
struct SIArgs {
  const int32_t* perm;
  const uint2* crec;   // 2D FP32 compact point records in sorted order, or null (sort.hip)
  const float* cmax;   // FP32 spreading: max |re|, |im| of the strengths per transform
  const int4* subprob;
  const int32_t* nsub;
  const void* coord[3];
  int64_t M;
  int nf[3];
  int bs[3];
  int nb[3];
  int tile[3];
  int64_t tile_cells;
  int64_t grid_cells;
  EsArgs<T> es;
  T dscale[3];  // -nf_d / (2 pi): chain factor of dz/dx
  EsPoly poly;  // polynomial form of the FP32 kernel at ns = 8 (on = 0 elsewhere)
};

template <typename T, typename CT, int D, int NS>
struct PointW {
  T w[D][NS];
  int i0[D];
};

template <typename T, typename CT, int D, int NS>
__device__ __forceinline__ void point_weights(const SIArgs<T>& a, int p, T (&w)[D][NS], int (&i0)[D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double x = (double)static_cast<const CT*>(a.coord[d])[p];
    const Folded f = fold_coord(x, a.nf[d]);
    i0[d] = support_start(f, NS / 2);
    const T t = (T)((double)i0[d] - f.xs);
    es_weights<T, NS>(t, a.es, w[d]);
  }
}

template <typename T, typename CT, int D, int NS>
__device__ __forceinline__ void point_weights_deriv(const SIArgs<T>& a, int p, T (&w)[D][NS], T (&dw)[D][NS],
                                                    int (&i0)[D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const double x = (double)static_cast<const CT*>(a.coord[d])[p];
    const Folded f = fold_coord(x, a.nf[d]);
    i0[d] = support_start(f, NS / 2);
    const T t = (T)((double)i0[d] - f.xs);
    es_weights_deriv<T, NS>(t, a.es, w[d], dw[d]);
  }
}

__device__ __forceinline__ int wrap_idx(int g, int n) {
  g %= n;
  return g < 0 ? g + n : g;
}

template <int D>
__device__ __forceinline__ void tile_origin(const int4& sd, const int (&bs)[3], const int (&nb)[3], int half,
                                            int (&org)[3]) {
  int rem = sd.x;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int bc = rem % nb[d];
    rem /= nb[d];
    org[d] = d < D ? bc * bs[d] - half : 0;
  }
}

template <int D>
__device__ __forceinline__ int64_t tile_cell_to_global(int64_t cell, const int (&tile)[3], const int (&org)[3],
                                                       const int (&nf)[3]) {
  const int l0 = (int)(cell % tile[0]);
  int64_t r = cell / tile[0];
  const int l1 = (int)(r % tile[1]);
  const int l2 = (int)(r / tile[1]);
  const int g0 = wrap_idx(org[0] + l0, nf[0]);
  if (D == 1) return g0;
  const int g1 = wrap_idx(org[1] + l1, nf[1]);
  if (D == 2) return (int64_t)g1 * nf[0] + g0;
  const int g2 = wrap_idx(org[2] + l2, nf[2]);
  return ((int64_t)g2 * nf[1] + g1) * nf[0] + g0;
}

// Rotated support walk for spreading.  Lane l visits support offset
// (k + rot_d) mod NS at step k of dimension d, with (rot_0, rot_1) distinct over
// the 64 lanes of a wave.  Lanes whose points share a support (dense regions,
// e.g. the centre of a radial trajectory) then update 64 distinct cells per
// instruction instead of serialising on one address; the weights are simply
// evaluated at the rotated offsets, so the rotation costs no extra work.
template <int NS>
__device__ __forceinline__ void lane_rotation(int (&rot)[3]) {
  const int lane = (int)(threadIdx.x & 63);
  rot[0] = lane % NS;
  rot[1] = (lane / NS) % NS;
  rot[2] = (rot[0] + rot[1]) % NS;
}

template <int NS>
__device__ __forceinline__ int rot_off(int k, int r) {
  int kk = k + r;
  return kk >= NS ? kk - NS : kk;
}

// rotated weights and support offsets (relative to i0) of one dimension
template <typename T, typename CT, int NS>
__device__ __forceinline__ void rot_weights(const SIArgs<T>& a, int d, int p, int rot, T (&w)[NS], int& i0, T& t) {
  const double x = (double)static_cast<const CT*>(a.coord[d])[p];
  const Folded f = fold_coord(x, a.nf[d]);
  i0 = support_start(f, NS / 2);
  t = (T)((double)i0 - f.xs);
#pragma unroll
  for (int k = 0; k < NS; ++k) w[k] = es_phi(t + (T)rot_off<NS>(k, rot), a.es.beta_log2e, a.es.c);
}

// ------------------------------------------------------------------ spread (tiled)
// LDS accumulation.  gfx950's LDS float atomics (ds_add_f32) run ~40x slower than
// its integer ones, so FP32 plans accumulate each tile in exact 64-bit fixed
// point: a contribution v becomes round(v * 2^s) through one FP64 FMA with the
// constant 1.5 * 2^52 (the rounded integer lands in the low mantissa bits), and
// is added with ds_add_u64.  s is chosen per subproblem from a bound B >= sum |Re c| +
// |Im c| of its points so that every partial sum stays below 2^50; the tile sum is
// then exact and independent of the order of the atomics.  FP64 plans use ds_add_f64.
template <typename T>
struct TileAcc;

template <>
struct TileAcc<float> {
  using Word = unsigned long long;
  static constexpr double kMagic = 6755399441055744.0;  // 1.5 * 2^52
  static __device__ __forceinline__ Word quant(double a, double b) {
    const double d = __builtin_fma(a, b, kMagic);
    return (Word)(__double_as_longlong(d) - 0x4338000000000000LL);
  }
};

// REC (2D FP32): points come from the sorted compact records (coalesced) instead of
// a gather and refold of the coordinates through perm.  A record word holds the
// support start in the tile (bits 25..31) and t + ns/2 + 1 in units of 2^-24.
template <typename T, int NS>
__device__ __forceinline__ void unpack_compact(unsigned w, int& l, T& t) {
  l = (int)(w >> 25);
  t = (T)((double)(w & 0x1FFFFFFu) * (1.0 / 16777216.0) - (double)(NS / 2 + 1));
}

template <typename T, typename CT, int D, int NS, bool REC = false>
__global__ void __launch_bounds__(256) k_spread_tiled(SIArgs<T> a, const cplx<T>* __restrict__ c,
                                                      cplx<T>* __restrict__ fw) {
  constexpr bool kFixed = std::is_same<T, float>::value;
  using Acc = typename std::conditional<kFixed, unsigned long long, double>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Acc* sre = reinterpret_cast<Acc*>(smem_raw);
  Acc* sim = sre + a.tile_cells;
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int tr = blockIdx.y;
  const int4 sd = a.subprob[sub];
  int org[3];
  tile_origin<D>(sd, a.bs, a.nb, NS / 2, org);
  for (int64_t i = threadIdx.x; i < a.tile_cells; i += blockDim.x) {
    sre[i] = (Acc)0;
    sim[i] = (Acc)0;
  }
  const cplx<T>* cT = c + (int64_t)tr * a.M;
  // fixed-point scale of this subproblem: every partial sum of a tile cell is below
  // B = sum_j (|re c_j| + |im c_j|) <= 2 n max|c| (max over the transform, one
  // coalesced reduction before the launch; weights are at most 1)
  double scale = 1.0, inv_scale = 1.0;
  bool finite = true;
  if constexpr (kFixed) {
    const double B = 2.0 * (double)sd.z * (double)a.cmax[tr];
    finite = isfinite(B);
    if (finite && B > 0.0) {
      int e;
      (void)frexp(B, &e);  // B < 2^e
      scale = ldexp(1.0, 50 - e);
      inv_scale = ldexp(1.0, e - 50);
    }
  }
  __syncthreads();
  const int t0 = a.tile[0], t1 = a.tile[1];
  int rot[3];
  lane_rotation<NS>(rot);
  for (int j = threadIdx.x; j < sd.z && finite; j += blockDim.x) {
    uint2 rc;
    if constexpr (REC) rc = a.crec[sd.y + j];
    const int p = a.perm[sd.y + j];
    if ((unsigned)p >= (unsigned)a.M) continue;
    const cplx<T> cv = cT[p];
    T w0[NS], w1[NS];
    int i00, i01 = 0;
    T tx, ty;
    int l0;
    if constexpr (REC) {
      unpack_compact<T, NS>(rc.x, l0, tx);
#pragma unroll
      for (int k = 0; k < NS; ++k) w0[k] = es_phi(tx + (T)rot_off<NS>(k, rot[0]), a.es.beta_log2e, a.es.c);
    } else {
      rot_weights<T, CT, NS>(a, 0, p, rot[0], w0, i00, tx);
      l0 = i00 - org[0];
    }
    int o0[NS];
    double wd[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      o0[k] = l0 + rot_off<NS>(k, rot[0]);
      wd[k] = (double)w0[k];
    }
    // add (vr, vi) * w to cell `idx` of the tile
    auto put = [&](int idx, double vr, double vi, int k) {
      if constexpr (kFixed) {
        atomicAdd(sre + idx, TileAcc<float>::quant(vr, wd[k]));
        atomicAdd(sim + idx, TileAcc<float>::quant(vi, wd[k]));
      } else {
        atomicAdd(sre + idx, vr * wd[k]);
        atomicAdd(sim + idx, vi * wd[k]);
      }
    };
    const double cre = (double)cv.re * scale, cim = (double)cv.im * scale;
    if constexpr (D == 1) {
#pragma unroll
      for (int k = 0; k < NS; ++k) put(o0[k], cre, cim, k);
    } else {
      int l1;
      if constexpr (REC) {
        unpack_compact<T, NS>(rc.y, l1, ty);
#pragma unroll
        for (int k = 0; k < NS; ++k) w1[k] = es_phi(ty + (T)rot_off<NS>(k, rot[1]), a.es.beta_log2e, a.es.c);
      } else {
        rot_weights<T, CT, NS>(a, 1, p, rot[1], w1, i01, ty);
        l1 = i01 - org[1];
      }
      if constexpr (D == 2) {
#pragma unroll
        for (int ky = 0; ky < NS; ++ky) {
          const double vr = cre * (double)w1[ky], vi = cim * (double)w1[ky];
          const int row = (l1 + rot_off<NS>(ky, rot[1])) * t0;
#pragma unroll
          for (int kx = 0; kx < NS; ++kx) put(row + o0[kx], vr, vi, kx);
        }
      } else {
        const double xz = (double)static_cast<const CT*>(a.coord[2])[p];
        const Folded fz = fold_coord(xz, a.nf[2]);
        const int i02 = support_start(fz, NS / 2);
        const T tz = (T)((double)i02 - fz.xs);
        const int l2 = i02 - org[2];
#pragma unroll 1
        for (int kz = 0; kz < NS; ++kz) {
          const int oz = rot_off<NS>(kz, rot[2]);
          const double wz = (double)es_phi(tz + (T)oz, a.es.beta_log2e, a.es.c);
          const double zr = cre * wz, zi = cim * wz;
          const int plane = (l2 + oz) * t1;
#pragma unroll
          for (int ky = 0; ky < NS; ++ky) {
            const double vr = zr * (double)w1[ky], vi = zi * (double)w1[ky];
            const int row = (plane + l1 + rot_off<NS>(ky, rot[1])) * t0;
#pragma unroll
            for (int kx = 0; kx < NS; ++kx) put(row + o0[kx], vr, vi, kx);
          }
        }
      }
    }
  }
  __syncthreads();
  cplx<T>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int64_t cell = threadIdx.x; cell < a.tile_cells; cell += blockDim.x) {
    T vr, vi;
    if constexpr (kFixed) {
      const long long ir = (long long)sre[cell], ii = (long long)sim[cell];
      if (finite && ir == 0 && ii == 0) continue;
      vr = finite ? (T)((double)ir * inv_scale) : (T)NAN;
      vi = finite ? (T)((double)ii * inv_scale) : (T)NAN;
    } else {
      vr = sre[cell];
      vi = sim[cell];
      if (vr == (T)0 && vi == (T)0) continue;
    }
    const int64_t gi = tile_cell_to_global<D>(cell, a.tile, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

// ------------------------------------------------------------------ spread (FP32, bank-spread)
// Same sums as k_spread_tiled<float, CT, D, NS> (the 64-bit fixed-point tile is exact, so
// the order in which a subproblem's points are added does not matter), with the points
// dealt to the lanes so that a wave's 64 atomics per step land on distinct LDS banks.
// Support cell (kz, ky, kx) of a point is at base + kz t0 t1 + ky t0 + kx with
// base = l2 t0 t1 + l1 t0 + l0, so at every step the lanes' addresses differ by their
// bases: bases covering the classes (base mod C) evenly give conflict-free atomics, where
// random bases serialise on shared banks.  Phase 1 stages every point in LDS (one compact
// record word per dimension, see unpack_compact, and its index) and counts the classes,
// phase 2 orders the staged points by class, phase 3 walks the class-ordered list
// transposed (lane q takes entry (q mod C) S + q / C, S = ceil(n / C)), so consecutive
// lanes take consecutive classes.
// C = 16 classes measured best on gfx950 (32 four-byte banks: 16 pairs for the 64-bit
// atomics; 32 and 64 classes were 2% and 9% slower on a 2D radial workload).
constexpr int kBankClasses = 16;

template <int D>
inline size_t banked_lds(int64_t tile_cells, int maxsub, bool packed = false) {
  return (size_t)tile_cells * (packed ? 8 : 16) +
         (size_t)maxsub * (D * sizeof(unsigned) + sizeof(int) + sizeof(unsigned short));
}

// Packed FP32 tiles (plans with eps >= kPackedTol): one 64-bit word per cell holds the real
// part in its high and the imaginary part in its low 32 bits, both signed fixed point, so a
// support cell takes one ds_add_u64 instead of two.  Word = re 2^32 + im (mod 2^64) adds
// exactly while every partial sum stays below 2^31 in magnitude: the scale is 2^(30 - e) with
// 2^e above 2 n max|c| (n points of the subproblem; frexp gives 2^e <= 4 n max|c|).  A unit is
// then at most 2^-18 max|c| for 1024 points, a rounding error near 4e-6 of the strengths, which
// the eps >= 1e-4 plans that use it do not see.
constexpr double kPackedTol = 1e-4;
constexpr int kPackedBits = 30;

// round(vr w) and round(vi w) packed, each rounded by one FP64 FMA with 1.5 2^52 (the integer
// lands in the low mantissa bits, two's complement): the low word of the real part's pattern
// is its integer mod 2^32, and the imaginary pattern minus the constant's is its integer.
__device__ __forceinline__ unsigned long long pack_re_im(double vr, double vi, double w) {
  const unsigned long long br = (unsigned long long)__double_as_longlong(__builtin_fma(vr, w, TileAcc<float>::kMagic));
  const unsigned long long bi = (unsigned long long)__double_as_longlong(__builtin_fma(vi, w, TileAcc<float>::kMagic));
  return (br << 32) + (bi - 0x4338000000000000ull);
}

__device__ __forceinline__ void unpack_re_im(unsigned long long w, long long& re, long long& im) {
  im = (long long)(int)(unsigned)(w & 0xFFFFFFFFull);
  re = (long long)(w - (unsigned long long)im) >> 32;
}

// compact record word of one coordinate (the layout of unpack_compact), tile-relative
template <int NS>
__device__ __forceinline__ unsigned pack_compact(double x, int nf, int org) {
  const Folded f = fold_coord(x, nf);
  const int i0 = support_start(f, NS / 2);
  const double u = ((double)i0 - f.xs + (double)(NS / 2 + 1)) * 16777216.0;  // 2^24
  const unsigned q = (unsigned)fmin(fmax(rint(u), 0.0), 33554431.0);      // 2^25 - 1
  return ((unsigned)(i0 - org) << 25) | q;
}

template <int D>
__device__ __forceinline__ int compact_base(const unsigned (&w)[D], int t0, int t01) {
  int b = (int)(w[0] >> 25);
  if constexpr (D >= 2) b += (int)(w[1] >> 25) * t0;
  if constexpr (D >= 3) b += (int)(w[2] >> 25) * t01;
  return b;
}

// Phases 1 and 2 of the bank-spread kernel: stage the n points of subproblem sd in LDS
// (word d of point j at sw[d maxsub + j], index at sp[j]) and write their class order into
// `order`.  ccnt[C] must be zero and visible to the block on entry; ends with a barrier.
template <typename CT, int D, int NS, bool REC, int C>
__device__ __forceinline__ void bank_order(const SIArgs<float>& a, const int4& sd, const int (&org)[3], int maxsub,
                                           unsigned* sw, int* sp, unsigned short* order, int* ccnt, int* cfill) {
  const int n = sd.z, t0 = a.tile[0], t01 = a.tile[0] * a.tile[1];
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const int p = a.perm[sd.y + j];
    unsigned w[D];
#pragma unroll
    for (int d = 0; d < D; ++d) w[d] = 0u;
    if constexpr (REC) {
      static_assert(D == 2, "compact records exist for 2D point sets");
      const uint2 r = a.crec[sd.y + j];
      w[0] = r.x;
      w[1] = r.y;
    } else if ((unsigned)p < (unsigned)a.M) {
#pragma unroll
      for (int d = 0; d < D; ++d) w[d] = pack_compact<NS>((double)static_cast<const CT*>(a.coord[d])[p], a.nf[d], org[d]);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) sw[d * maxsub + j] = w[d];
    sp[j] = p;
    atomicAdd(&ccnt[compact_base<D>(w, t0, t01) & (C - 1)], 1);
  }
  __syncthreads();
  if (threadIdx.x < 64) {  // exclusive scan of the class counts (one wave)
    const int lane = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < C; base += 64) {
      const int v = base + lane < C ? ccnt[base + lane] : 0;
      int incl = v;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
      }
      if (base + lane < C) cfill[base + lane] = carry + incl - v;
      carry += __shfl(incl, 63, 64);
    }
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    unsigned w[D];
#pragma unroll
    for (int d = 0; d < D; ++d) w[d] = sw[d * maxsub + j];
    const int slot = atomicAdd(&cfill[compact_base<D>(w, t0, t01) & (C - 1)], 1);
    order[slot] = (unsigned short)j;
  }
  __syncthreads();
}

// One block of the bank-spread kernel: subproblem sub of transform tr.  PK: packed tile.
template <typename CT, int D, int NS, bool REC, int C, bool PK = false>
__device__ __forceinline__ void spread_banked_block(const SIArgs<float>& a, int maxsub, int sub, int tr,
                                                    const cplx<float>* __restrict__ c, cplx<float>* __restrict__ fw) {
  using Word = unsigned long long;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  Word* sre = reinterpret_cast<Word*>(smem_raw);
  Word* sim = sre + (PK ? 0 : a.tile_cells);  // unused when packed
  unsigned* sw = reinterpret_cast<unsigned*>(sre + (PK ? 1 : 2) * a.tile_cells);
  int* sp = reinterpret_cast<int*>(sw + (size_t)D * maxsub);
  unsigned short* order = reinterpret_cast<unsigned short*>(sp + maxsub);
  __shared__ int ccnt[C], cfill[C];
  if (sub >= *a.nsub) return;
  const int4 sd = a.subprob[sub];
  const int n = sd.z;
  int org[3];
  tile_origin<D>(sd, a.bs, a.nb, NS / 2, org);
  for (int64_t i = threadIdx.x; i < a.tile_cells; i += blockDim.x) {
    sre[i] = 0;
    if constexpr (!PK) sim[i] = 0;
  }
  if (threadIdx.x < C) ccnt[threadIdx.x] = 0;
  // fixed-point scale, as in k_spread_tiled (packed: 30 instead of 50 bits below 2^e)
  constexpr int kBits = PK ? kPackedBits : 50;
  const double B = 2.0 * (double)n * (double)a.cmax[tr];
  const bool finite = isfinite(B);
  double scale = 1.0, inv_scale = 1.0;
  if (finite && B > 0.0) {
    int e;
    (void)frexp(B, &e);
    scale = ldexp(1.0, kBits - e);
    inv_scale = ldexp(1.0, e - kBits);
  }
  __syncthreads();
  bank_order<CT, D, NS, REC, C>(a, sd, org, maxsub, sw, sp, order, ccnt, cfill);
  const int t0 = a.tile[0], t01 = a.tile[0] * a.tile[1];
  // ---- phase 3: accumulate, lanes dealt across the classes
  const cplx<float>* cT = c + (int64_t)tr * a.M;
  const int S = (n + C - 1) / C;
  const int nq = S * C;
  for (int q = threadIdx.x; q < nq && finite; q += blockDim.x) {
    const int pos = (q % C) * S + q / C;
    if (pos >= n) continue;
    const int j = order[pos];
    const int p = sp[j];
    if ((unsigned)p >= (unsigned)a.M) continue;
    const cplx<float> cv = cT[p];
    int l[3] = {0, 0, 0};
    float t[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int d = 0; d < D; ++d) unpack_compact<float, NS>(sw[d * maxsub + j], l[d], t[d]);
    const int base = l[0] + l[1] * t0 + l[2] * t01;
    double wd[NS];
    float w1[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      wd[k] = (double)es_phi(t[0] + (float)k, a.es.beta_log2e, a.es.c);
      w1[k] = es_phi(t[1] + (float)k, a.es.beta_log2e, a.es.c);
    }
    const double cre = (double)cv.re * scale, cim = (double)cv.im * scale;
    const int kzn = D == 3 ? NS : 1;
#pragma unroll 1
    for (int kz = 0; kz < kzn; ++kz) {
      double zr = cre, zi = cim;
      if constexpr (D == 3) {
        const double wz = (double)es_phi(t[2] + (float)kz, a.es.beta_log2e, a.es.c);
        zr *= wz;
        zi *= wz;
      }
      const int plane = base + kz * t01;
#pragma unroll
      for (int ky = 0; ky < NS; ++ky) {
        const double vr = zr * (double)w1[ky], vi = zi * (double)w1[ky];
        const int row = plane + ky * t0;
#pragma unroll
        for (int kx = 0; kx < NS; ++kx) {
          if constexpr (PK) {
            atomicAdd(sre + row + kx, pack_re_im(vr, vi, wd[kx]));
          } else {
            atomicAdd(sre + row + kx, TileAcc<float>::quant(vr, wd[kx]));
            atomicAdd(sim + row + kx, TileAcc<float>::quant(vi, wd[kx]));
          }
        }
      }
    }
  }
  __syncthreads();
  cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int64_t cell = threadIdx.x; cell < a.tile_cells; cell += blockDim.x) {
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
    const int64_t gi = tile_cell_to_global<D>(cell, a.tile, org, a.nf);
    atomic_add_global(&fT[gi].re, vr);
    atomic_add_global(&fT[gi].im, vi);
  }
}

template <typename CT, int D, int NS, bool REC, int C, bool PK = false>
__global__ void __launch_bounds__(1024) k_spread_banked(SIArgs<float> a, int maxsub, const cplx<float>* __restrict__ c,
                                                        cplx<float>* __restrict__ fw) {
  spread_banked_block<CT, D, NS, REC, C, PK>(a, maxsub, (int)blockIdx.x, (int)blockIdx.y, c, fw);
}

// Several point sets (plans of one kernel size and dimension, any points and grids) in one
// launch: plan p owns blocks [first[p], first[p+1]) of the x dimension.
constexpr int kMaxMany = kMaxManyPlans;

struct SpreadMany {
  SIArgs<float> a[kMaxMany];
  const cplx<float>* c[kMaxMany];
  cplx<float>* fw[kMaxMany];
  int first[kMaxMany + 1];
  int count, maxsub;
};
static_assert(sizeof(SpreadMany) <= 4096, "kernel argument block too large");

__device__ __forceinline__ int many_plan(const int* first, int count, int b) {
  int p = 0;
  for (int q = 1; q < count; ++q)
    if (first[q] <= b) p = q;
  return p;
}

template <typename CT, int D, int NS, bool REC, int C>
__global__ void __launch_bounds__(1024) k_spread_banked_many(SpreadMany m) {
  const int p = many_plan(m.first, m.count, (int)blockIdx.x);
  spread_banked_block<CT, D, NS, REC, C>(m.a[p], m.maxsub, (int)blockIdx.x - m.first[p], (int)blockIdx.y, m.c[p],
                                         m.fw[p]);
}

// ------------------------------------------------------------------ spread (global)
template <typename T, typename CT, int D, int NS>
__global__ void __launch_bounds__(256) k_spread_global(SIArgs<T> a, const cplx<T>* __restrict__ c,
                                                       cplx<T>* __restrict__ fw) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.M) return;
  const int tr = blockIdx.y;
  const int p = a.perm[j];
  if ((unsigned)p >= (unsigned)a.M) return;
  const cplx<T> cv = c[(int64_t)tr * a.M + p];
  int rot[3];
  lane_rotation<NS>(rot);
  T w0[NS], w1[NS];
  int i00, i01 = 0;
  T tx, ty;
  rot_weights<T, CT, NS>(a, 0, p, rot[0], w0, i00, tx);
  cplx<T>* fT = fw + (int64_t)tr * a.grid_cells;
  int gx[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) gx[k] = wrap_idx(i00 + rot_off<NS>(k, rot[0]), a.nf[0]);
  if constexpr (D == 1) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      atomic_add_global(&fT[gx[k]].re, cv.re * w0[k]);
      atomic_add_global(&fT[gx[k]].im, cv.im * w0[k]);
    }
  } else {
    rot_weights<T, CT, NS>(a, 1, p, rot[1], w1, i01, ty);
    int i02 = 0;
    T tz = 0;
    if constexpr (D == 3) {
      const double xz = (double)static_cast<const CT*>(a.coord[2])[p];
      const Folded fz = fold_coord(xz, a.nf[2]);
      i02 = support_start(fz, NS / 2);
      tz = (T)((double)i02 - fz.xs);
    }
    const int kzn = D == 3 ? NS : 1;
#pragma unroll 1
    for (int kz = 0; kz < kzn; ++kz) {
      T zr = cv.re, zi = cv.im;
      int64_t zoff = 0;
      if constexpr (D == 3) {
        const int oz = rot_off<NS>(kz, rot[2]);
        const T wz = es_phi(tz + (T)oz, a.es.beta_log2e, a.es.c);
        zr *= wz;
        zi *= wz;
        zoff = (int64_t)wrap_idx(i02 + oz, a.nf[2]) * a.nf[1];
      }
#pragma unroll
      for (int ky = 0; ky < NS; ++ky) {
        const T vr = zr * w1[ky], vi = zi * w1[ky];
        const int64_t row = (zoff + wrap_idx(i01 + rot_off<NS>(ky, rot[1]), a.nf[1])) * a.nf[0];
#pragma unroll
        for (int kx = 0; kx < NS; ++kx) {
          atomic_add_global(&fT[row + gx[kx]].re, vr * w0[kx]);
          atomic_add_global(&fT[row + gx[kx]].im, vi * w0[kx]);
        }
      }
    }
  }
}

// Global-memory 2D/3D spread, one wave per point (grids whose tiles do not fit the LDS: the
// sparse FP64 grids of type 3 and high-accuracy plans).  The lanes first evaluate the point's
// D NS kernel weights at once (lane q < D NS: weight q mod NS of dimension q / NS), then every
// atomic instruction covers R = 64 / (2 NS) support rows of one z slice, lane (r, 2 kx + part)
// adding part (re, im) of cell kx of row r: the wave's addresses are R runs of 2 NS adjacent
// words, a few cache lines per instruction where one point per thread touches 64 unrelated
// ones.  Weights and support starts are those of k_spread_global (same fold, same es_phi);
// the product is taken as c wz wx wy.
template <typename T, typename CT, int D, int NS>
__global__ void __launch_bounds__(256) k_spread_global_wave(SIArgs<T> a, const cplx<T>* __restrict__ c,
                                                            cplx<T>* __restrict__ fw) {
  constexpr int RL = 2 * NS;  // lanes per support row
  constexpr int R = 64 / RL;  // support rows per instruction
  static_assert((D == 2 || D == 3) && D * NS <= 64 && R >= 1, "one wave holds the D NS weights and a support row");
  const int lane = (int)(threadIdx.x & 63);
  const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // sorted slot of this wave's point
  if (j >= a.M) return;                                                     // uniform over the wave
  const int tr = blockIdx.y;
  const int p = a.perm[j];
  if ((unsigned)p >= (unsigned)a.M) return;
  const int dq = lane / NS, kq = lane - dq * NS;
  T wq = (T)0;
  int i0q = 0;
  if (dq < D) {
    const double x = (double)static_cast<const CT*>(a.coord[dq])[p];
    const Folded f = fold_coord(x, a.nf[dq]);
    i0q = support_start(f, NS / 2);
    wq = es_phi((T)((double)i0q - f.xs) + (T)kq, a.es.beta_log2e, a.es.c);
  }
  const int i0x = __shfl(i0q, 0, 64), i0y = __shfl(i0q, NS, 64), i0z = __shfl(i0q, D == 3 ? 2 * NS : 0, 64);
  const int rr = lane / RL, q = lane - rr * RL, kx = q >> 1, part = q & 1;
  const T wx = __shfl(wq, kx < NS ? kx : 0, 64);
  const cplx<T> cv = c[(int64_t)tr * a.M + p];
  const T cx = (part ? cv.im : cv.re) * wx;
  const int64_t gx = wrap_idx(i0x + kx, a.nf[0]);
  T* fT = reinterpret_cast<T*>(fw + (int64_t)tr * a.grid_cells);
#pragma unroll 1
  for (int kz = 0; kz < (D == 3 ? NS : 1); ++kz) {
    const T wz = D == 3 ? __shfl(wq, 2 * NS + kz, 64) : (T)1;
    const int64_t zoff = D == 3 ? (int64_t)wrap_idx(i0z + kz, a.nf[2]) * a.nf[1] : 0;
    const T cz = cx * wz;
#pragma unroll
    for (int r0 = 0; r0 < NS; r0 += R) {
      const int ky = r0 + rr;
      const T wy = __shfl(wq, NS + (ky < NS ? ky : NS - 1), 64);
      if (rr < R && ky < NS) {
        const int64_t cell = (zoff + wrap_idx(i0y + ky, a.nf[1])) * a.nf[0] + gx;
        atomic_add_global(&fT[2 * cell + part], cz * wy);
      }
    }
  }
}

// ------------------------------------------------------------------ gather helpers
// Accumulates sum over the support of G[cell] * wx[kx] * wy[ky] * wz[kz] for
// value weights, and optionally the per-dimension derivative sums.
template <typename T, int D, int NS, bool DERIV, typename Fetch>
__device__ __forceinline__ void gather_point(const T (&w)[D][NS], const T (&dw)[D][NS], Fetch fetch, cplx<T>& s0,
                                             cplx<T> (&sd)[D]) {
  s0 = {0, 0};
#pragma unroll
  for (int d = 0; d < D; ++d) sd[d] = {0, 0};
  if constexpr (D == 1) {
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      const cplx<T> v = fetch(0, 0, k);
      s0.re += v.re * w[0][k];
      s0.im += v.im * w[0][k];
      if constexpr (DERIV) {
        sd[0].re += v.re * dw[0][k];
        sd[0].im += v.im * dw[0][k];
      }
    }
  } else {
    const int kzn = D == 3 ? NS : 1;
#pragma unroll 1
    for (int kz = 0; kz < kzn; ++kz) {
      cplx<T> p0 = {0, 0}, px = {0, 0}, py = {0, 0};
#pragma unroll
      for (int ky = 0; ky < NS; ++ky) {
        cplx<T> r0 = {0, 0}, rx = {0, 0};
#pragma unroll
        for (int kx = 0; kx < NS; ++kx) {
          const cplx<T> v = fetch(kz, ky, kx);
          r0.re += v.re * w[0][kx];
          r0.im += v.im * w[0][kx];
          if constexpr (DERIV) {
            rx.re += v.re * dw[0][kx];
            rx.im += v.im * dw[0][kx];
          }
        }
        p0.re += r0.re * w[1][ky];
        p0.im += r0.im * w[1][ky];
        if constexpr (DERIV) {
          px.re += rx.re * w[1][ky];
          px.im += rx.im * w[1][ky];
          py.re += r0.re * dw[1][ky];
          py.im += r0.im * dw[1][ky];
        }
      }
      if constexpr (D == 2) {
        s0 = p0;
        if constexpr (DERIV) {
          sd[0] = px;
          sd[1] = py;
        }
      } else {
        const T wz = w[D - 1][kz];
        s0.re += p0.re * wz;
        s0.im += p0.im * wz;
        if constexpr (DERIV) {
          const T dwz = dw[D - 1][kz];
          sd[0].re += px.re * wz;
          sd[0].im += px.im * wz;
          sd[1].re += py.re * wz;
          sd[1].im += py.im * wz;
          sd[D - 1].re += p0.re * dwz;
          sd[D - 1].im += p0.im * dwz;
        }
      }
    }
  }
}

// Output stage shared by interp and the gradient gather.
template <typename T, int D, bool DERIV>
struct GatherOut {
  cplx<T>* val;           // [n_trans][M] or nullptr
  const cplx<T>* P;       // [n_trans][M] pointwise factor for coordinate gradients
  T* gcoord[3];           // [M] each, accumulated, or nullptr
};

template <typename T, int D, bool DERIV>
__device__ __forceinline__ void emit_point(const SIArgs<T>& a, const GatherOut<T, D, DERIV>& o, int tr, int p,
                                           const cplx<T>& s0, const cplx<T> (&sd)[D]) {
  const int64_t ip = (int64_t)tr * a.M + p;
  if (o.val) o.val[ip] = s0;
  if constexpr (DERIV) {
    const cplx<T> pv = o.P[ip];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (o.gcoord[d]) {
        // Re(conj(P) * S) with the dz/dx chain factor
        const T g = (pv.re * sd[d].re + pv.im * sd[d].im) * a.dscale[d];
        atomic_add_global(o.gcoord[d] + p, g);
      }
    }
  }
}

// ------------------------------------------------------------------ interp / grad (tiled)
// REC (2D FP32, values only): points come from the sorted point records.
template <typename T, typename CT, int D, int NS, bool DERIV, bool REC>
__device__ __forceinline__ void gather_tiled_block(const SIArgs<T>& a, const cplx<T>* __restrict__ fw,
                                                   const GatherOut<T, D, DERIV>& o, int sub, int tr) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<T>* tile = reinterpret_cast<cplx<T>*>(smem_raw);
  if (sub >= *a.nsub) return;
  const int4 sd4 = a.subprob[sub];
  int org[3];
  tile_origin<D>(sd4, a.bs, a.nb, NS / 2, org);
  const cplx<T>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int64_t cell = threadIdx.x; cell < a.tile_cells; cell += blockDim.x)
    tile[cell] = fT[tile_cell_to_global<D>(cell, a.tile, org, a.nf)];
  __syncthreads();
  const int t0 = a.tile[0], t1 = a.tile[1];
  for (int j = threadIdx.x; j < sd4.z; j += blockDim.x) {
    T w[D][NS], dw[D][NS];
    int l[3] = {0, 0, 0};
    int p;
    if constexpr (REC) {
      const uint2 rc = a.crec[sd4.y + j];
      p = a.perm[sd4.y + j];
      T tx, ty;
      unpack_compact<T, NS>(rc.x, l[0], tx);
      unpack_compact<T, NS>(rc.y, l[1], ty);
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        w[0][k] = es_phi(tx + (T)k, a.es.beta_log2e, a.es.c);
        w[D - 1][k] = es_phi(ty + (T)k, a.es.beta_log2e, a.es.c);
      }
    } else {
      p = a.perm[sd4.y + j];
      if ((unsigned)p >= (unsigned)a.M) continue;
      int i0[D];
      if constexpr (DERIV)
        point_weights_deriv<T, CT, D, NS>(a, p, w, dw, i0);
      else
        point_weights<T, CT, D, NS>(a, p, w, i0);
#pragma unroll
      for (int d = 0; d < D; ++d) l[d] = i0[d] - org[d];
    }
    auto fetch = [&](int kz, int ky, int kx) -> cplx<T> {
      return tile[((int64_t)(l[2] + kz) * t1 + (l[1] + ky)) * t0 + l[0] + kx];
    };
    cplx<T> s0, sdd[D];
    gather_point<T, D, NS, DERIV>(w, dw, fetch, s0, sdd);
    emit_point<T, D, DERIV>(a, o, tr, p, s0, sdd);
  }
}

template <typename T, typename CT, int D, int NS, bool DERIV, bool REC = false>
__global__ void __launch_bounds__(256) k_gather_tiled(SIArgs<T> a, const cplx<T>* __restrict__ fw,
                                                      GatherOut<T, D, DERIV> o) {
  gather_tiled_block<T, CT, D, NS, DERIV, REC>(a, fw, o, (int)blockIdx.x, (int)blockIdx.y);
}

// FP32 interpolation with 16-byte tile reads: a support row of NS cells starting at l0 is read
// as NS/2 + 1 aligned cell pairs from l0 rounded down to even (tile rows of an even number of
// cells keep every pair 16-byte aligned), with the x weights shifted by one cell when l0 is
// odd and zero on the extra cell; two zero cells after the tile cover the last row's overrun.
// REC (2D): points from the compact records; otherwise folded from the coordinates.
// POLY: kernel weights from the polynomial ES kernel (ns = 8, common.hpp).
template <typename CT, int D, int NS, bool REC, bool POLY = false>
__global__ void __launch_bounds__(256) k_gather_wide(SIArgs<float> a, const cplx<float>* __restrict__ fw,
                                                     cplx<float>* __restrict__ out) {
  constexpr int NC = NS + 2;  // cells read per row
  constexpr int NQ = NC / 2;  // 16-byte reads per row
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<float>* tile = reinterpret_cast<cplx<float>*>(smem_raw);
  const int sub = blockIdx.x, tr = blockIdx.y;
  if (sub >= *a.nsub) return;
  const int4 sd4 = a.subprob[sub];
  int org[3];
  tile_origin<D>(sd4, a.bs, a.nb, NS / 2, org);
  const cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int64_t cell = threadIdx.x; cell < a.tile_cells; cell += blockDim.x)
    tile[cell] = fT[tile_cell_to_global<D>(cell, a.tile, org, a.nf)];
  if (threadIdx.x < 2) tile[a.tile_cells + threadIdx.x] = cplx<float>{0.0f, 0.0f};
  __syncthreads();
  const int t0 = a.tile[0], t1 = a.tile[1];
  const int hq = t0 >> 1;  // pairs per tile row
  for (int j = threadIdx.x; j < sd4.z; j += blockDim.x) {
    const int p = a.perm[sd4.y + j];
    int l[3] = {0, 0, 0};
    float t[3] = {0.0f, 0.0f, 0.0f};  // t = i0 - xs, or (POLY) v = 2 (t + NS/2) - 1
    if constexpr (REC) {
      const uint2 rc = a.crec[sd4.y + j];
      unpack_compact<float, NS>(rc.x, l[0], t[0]);
      unpack_compact<float, NS>(rc.y, l[1], t[1]);
      if constexpr (POLY) {
        t[0] = es_poly_v(rc.x);
        t[1] = es_poly_v(rc.y);
      }
    } else {
      if ((unsigned)p >= (unsigned)a.M) continue;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const Folded f = fold_coord((double)static_cast<const CT*>(a.coord[d])[p], a.nf[d]);
        const int i0 = support_start(f, NS / 2);
        l[d] = i0 - org[d];
        t[d] = POLY ? (float)(2.0 * ((double)i0 - f.xs) + (double)(NS - 1)) : (float)((double)i0 - f.xs);
      }
    }
    // t + k lies in [-NS/2, NS/2] (t in [-NS/2, 1 - NS/2)): no range select needed
    float wx[NS], wy[NS], wz[NS];
    if constexpr (POLY) {
      es_poly_weights<NS>(t[0], a.poly, wx);
      es_poly_weights<NS>(t[1], a.poly, wy);
      if constexpr (D == 3) {
        es_poly_weights<NS>(t[2], a.poly, wz);
      } else {
#pragma unroll
        for (int k = 0; k < NS; ++k) wz[k] = 1.0f;
      }
    } else {
#pragma unroll
      for (int k = 0; k < NS; ++k) {
        wx[k] = es_phi_in(t[0] + (float)k, a.es.beta_log2e, a.es.c);
        wy[k] = es_phi_in(t[1] + (float)k, a.es.beta_log2e, a.es.c);
        wz[k] = D == 3 ? es_phi_in(t[2] + (float)k, a.es.beta_log2e, a.es.c) : 1.0f;
      }
    }
    const bool odd = (l[0] & 1) != 0;
    float ws[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const float we = k < NS ? wx[k] : 0.0f;
      const float wo = (k >= 1 && k <= NS) ? wx[k - 1] : 0.0f;
      ws[k] = odd ? wo : we;
    }
    const float4* base = reinterpret_cast<const float4*>(
        tile + ((int64_t)l[2] * t1 + l[1]) * t0 + (l[0] - (odd ? 1 : 0)));
    float sr = 0.0f, si = 0.0f;
    const int kzn = D == 3 ? NS : 1;
#pragma unroll 1
    for (int kz = 0; kz < kzn; ++kz) {
      const float4* plane = base + (int64_t)kz * t1 * hq;
      float pr = 0.0f, pi = 0.0f;
#pragma unroll
      for (int ky = 0; ky < NS; ++ky) {
        const float4* q = plane + (int64_t)ky * hq;
        float rr = 0.0f, ri = 0.0f;
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
          const float4 v = q[u];
          rr = fmaf(v.x, ws[2 * u], rr);
          ri = fmaf(v.y, ws[2 * u], ri);
          rr = fmaf(v.z, ws[2 * u + 1], rr);
          ri = fmaf(v.w, ws[2 * u + 1], ri);
        }
        pr = fmaf(rr, wy[ky], pr);
        pi = fmaf(ri, wy[ky], pi);
      }
      sr = fmaf(pr, wz[kz], sr);
      si = fmaf(pi, wz[kz], si);
    }
    if ((unsigned)p < (unsigned)a.M) out[(int64_t)tr * a.M + p] = cplx<float>{sr, si};
  }
}

// The coordinate-gradient gather (values, derivative sums; emit_point applies P and the chain
// factor) with the 16-byte pair reads of k_gather_wide: FP32, points folded from coordinates.
template <typename CT, int D, int NS>
__global__ void __launch_bounds__(256) k_gather_wide_deriv(SIArgs<float> a, const cplx<float>* __restrict__ fw,
                                                           GatherOut<float, D, true> o) {
  constexpr int NC = NS + 2;
  constexpr int NQ = NC / 2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  cplx<float>* tile = reinterpret_cast<cplx<float>*>(smem_raw);
  const int sub = blockIdx.x, tr = blockIdx.y;
  if (sub >= *a.nsub) return;
  const int4 sd4 = a.subprob[sub];
  int org[3];
  tile_origin<D>(sd4, a.bs, a.nb, NS / 2, org);
  const cplx<float>* fT = fw + (int64_t)tr * a.grid_cells;
  for (int64_t cell = threadIdx.x; cell < a.tile_cells; cell += blockDim.x)
    tile[cell] = fT[tile_cell_to_global<D>(cell, a.tile, org, a.nf)];
  if (threadIdx.x < 2) tile[a.tile_cells + threadIdx.x] = cplx<float>{0.0f, 0.0f};
  __syncthreads();
  const int t0 = a.tile[0], t1 = a.tile[1];
  const int hq = t0 >> 1;
  for (int j = threadIdx.x; j < sd4.z; j += blockDim.x) {
    const int p = a.perm[sd4.y + j];
    if ((unsigned)p >= (unsigned)a.M) continue;
    float w[D][NS], dw[D][NS];
    int i0[D];
    point_weights_deriv<float, CT, D, NS>(a, p, w, dw, i0);
    int l[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) l[d] = i0[d] - org[d];
    const bool odd = (l[0] & 1) != 0;
    float ws[NC], dws[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const bool ine = k < NS, ino = k >= 1 && k <= NS;
      ws[k] = odd ? (ino ? w[0][k - 1] : 0.0f) : (ine ? w[0][k] : 0.0f);
      dws[k] = odd ? (ino ? dw[0][k - 1] : 0.0f) : (ine ? dw[0][k] : 0.0f);
    }
    const float4* base = reinterpret_cast<const float4*>(
        tile + ((int64_t)l[2] * t1 + l[1]) * t0 + (l[0] - (odd ? 1 : 0)));
    cplx<float> s0 = {0.0f, 0.0f}, sd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) sd[d] = {0.0f, 0.0f};
    const int kzn = D == 3 ? NS : 1;
#pragma unroll 1
    for (int kz = 0; kz < kzn; ++kz) {
      const float4* plane = base + (int64_t)kz * t1 * hq;
      cplx<float> p0 = {0.0f, 0.0f}, px = {0.0f, 0.0f}, py = {0.0f, 0.0f};
#pragma unroll
      for (int ky = 0; ky < NS; ++ky) {
        const float4* q = plane + (int64_t)ky * hq;
        cplx<float> r0 = {0.0f, 0.0f}, rx = {0.0f, 0.0f};
#pragma unroll
        for (int u = 0; u < NQ; ++u) {
          const float4 v = q[u];
          r0.re = fmaf(v.x, ws[2 * u], fmaf(v.z, ws[2 * u + 1], r0.re));
          r0.im = fmaf(v.y, ws[2 * u], fmaf(v.w, ws[2 * u + 1], r0.im));
          rx.re = fmaf(v.x, dws[2 * u], fmaf(v.z, dws[2 * u + 1], rx.re));
          rx.im = fmaf(v.y, dws[2 * u], fmaf(v.w, dws[2 * u + 1], rx.im));
        }
        const float wy = w[1][ky], dwy = dw[1][ky];
        p0.re = fmaf(r0.re, wy, p0.re);
        p0.im = fmaf(r0.im, wy, p0.im);
        px.re = fmaf(rx.re, wy, px.re);
        px.im = fmaf(rx.im, wy, px.im);
        py.re = fmaf(r0.re, dwy, py.re);
        py.im = fmaf(r0.im, dwy, py.im);
      }
      if constexpr (D == 2) {
        s0 = p0;
        sd[0] = px;
        sd[1] = py;
      } else {
        const float wz = w[D - 1][kz], dwz = dw[D - 1][kz];
        s0.re = fmaf(p0.re, wz, s0.re);
        s0.im = fmaf(p0.im, wz, s0.im);
        sd[0].re = fmaf(px.re, wz, sd[0].re);
        sd[0].im = fmaf(px.im, wz, sd[0].im);
        sd[1].re = fmaf(py.re, wz, sd[1].re);
        sd[1].im = fmaf(py.im, wz, sd[1].im);
        sd[D - 1].re = fmaf(p0.re, dwz, sd[D - 1].re);
        sd[D - 1].im = fmaf(p0.im, dwz, sd[D - 1].im);
      }
    }
    emit_point<float, D, true>(a, o, tr, p, s0, sd);
  }
}

// Values-only FP32 interpolation of several point sets in one launch (see SpreadMany).
struct GatherMany {
  SIArgs<float> a[kMaxMany];
  const cplx<float>* fw[kMaxMany];
  cplx<float>* out[kMaxMany];
  int first[kMaxMany + 1];
  int count;
};
static_assert(sizeof(GatherMany) <= 4096, "kernel argument block too large");

template <typename CT, int D, int NS, bool REC>
__global__ void __launch_bounds__(256) k_gather_tiled_many(GatherMany m) {
  const int p = many_plan(m.first, m.count, (int)blockIdx.x);
  GatherOut<float, D, false> o;
  o.val = m.out[p];
  o.P = nullptr;
  o.gcoord[0] = o.gcoord[1] = o.gcoord[2] = nullptr;
  gather_tiled_block<float, CT, D, NS, false, REC>(m.a[p], m.fw[p], o, (int)blockIdx.x - m.first[p],
                                                   (int)blockIdx.y);
}

template <typename T, typename CT, int D, int NS, bool DERIV>
__global__ void __launch_bounds__(256) k_gather_global(SIArgs<T> a, const cplx<T>* __restrict__ fw,
                                                       GatherOut<T, D, DERIV> o) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= a.M) return;
  const int tr = blockIdx.y;
  const int p = a.perm[j];
  if ((unsigned)p >= (unsigned)a.M) return;
  T w[D][NS], dw[D][NS];
  int i0[D];
  if constexpr (DERIV)
    point_weights_deriv<T, CT, D, NS>(a, p, w, dw, i0);
  else
    point_weights<T, CT, D, NS>(a, p, w, i0);
  const cplx<T>* fT = fw + (int64_t)tr * a.grid_cells;
  auto fetch = [&](int kz, int ky, int kx) -> cplx<T> {
    int64_t idx = wrap_idx(i0[0] + kx, a.nf[0]);
    if (D >= 2) idx += (int64_t)wrap_idx(i0[D >= 2 ? 1 : 0] + ky, a.nf[1]) * a.nf[0];
    if (D == 3) idx += (int64_t)wrap_idx(i0[D - 1] + kz, a.nf[2]) * a.nf[1] * a.nf[0];
    return fT[idx];
  };
  cplx<T> s0, sdd[D];
  gather_point<T, D, NS, DERIV>(w, dw, fetch, s0, sdd);
  emit_point<T, D, DERIV>(a, o, tr, p, s0, sdd);
}

// Global-memory 2D/3D interpolation, one wave per point (the transpose of
// k_spread_global_wave): after the weights, every load instruction covers R = 64 / (2 NS)
// support rows of one z slice, lane (r, 2 kx + part) reading part (re, im) of cell kx of row r
// (R runs of 2 NS adjacent words), and multiplies it by its wx wy wz; the real and imaginary
// sums end on lanes 0 and 1 after a butterfly over the lanes of equal parity.
template <typename T, typename CT, int D, int NS>
__global__ void __launch_bounds__(256) k_gather_global_wave(SIArgs<T> a, const cplx<T>* __restrict__ fw,
                                                            cplx<T>* __restrict__ out) {
  constexpr int RL = 2 * NS;  // lanes per support row
  constexpr int R = 64 / RL;  // support rows per instruction
  static_assert((D == 2 || D == 3) && D * NS <= 64 && R >= 1, "one wave holds the D NS weights and a support row");
  const int lane = (int)(threadIdx.x & 63);
  const int64_t j = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;  // sorted slot of this wave's point
  if (j >= a.M) return;                                                     // uniform over the wave
  const int tr = blockIdx.y;
  const int p = a.perm[j];
  if ((unsigned)p >= (unsigned)a.M) return;
  const int dq = lane / NS, kq = lane - dq * NS;
  T wq = (T)0;
  int i0q = 0;
  if (dq < D) {
    const double x = (double)static_cast<const CT*>(a.coord[dq])[p];
    const Folded f = fold_coord(x, a.nf[dq]);
    i0q = support_start(f, NS / 2);
    wq = es_phi((T)((double)i0q - f.xs) + (T)kq, a.es.beta_log2e, a.es.c);
  }
  const int i0x = __shfl(i0q, 0, 64), i0y = __shfl(i0q, NS, 64), i0z = __shfl(i0q, D == 3 ? 2 * NS : 0, 64);
  const int rr = lane / RL, q = lane - rr * RL, kx = q >> 1, part = q & 1;
  const T wx = __shfl(wq, kx < NS ? kx : 0, 64);
  const int64_t gx = wrap_idx(i0x + kx, a.nf[0]);
  const T* fT = reinterpret_cast<const T*>(fw + (int64_t)tr * a.grid_cells);
  T acc = (T)0;
#pragma unroll 1
  for (int kz = 0; kz < (D == 3 ? NS : 1); ++kz) {
    const T wz = D == 3 ? __shfl(wq, 2 * NS + kz, 64) : (T)1;
    const int64_t zoff = D == 3 ? (int64_t)wrap_idx(i0z + kz, a.nf[2]) * a.nf[1] : 0;
    T accz = (T)0;
#pragma unroll
    for (int r0 = 0; r0 < NS; r0 += R) {
      const int ky = r0 + rr;
      const T wy = __shfl(wq, NS + (ky < NS ? ky : NS - 1), 64);
      if (rr < R && ky < NS) {
        const int64_t cell = (zoff + wrap_idx(i0y + ky, a.nf[1])) * a.nf[0] + gx;
        accz += fT[2 * cell + part] * wy;
      }
    }
    acc += accz * wz;
  }
  acc *= wx;
  // lanes of equal parity (2 kx + part keeps part = lane & 1, RL being even) hold one part
#pragma unroll
  for (int off = 2; off < 64; off <<= 1) acc += __shfl_xor(acc, off, 64);
  const T im = __shfl(acc, 1, 64);
  if (lane == 0) out[(int64_t)tr * a.M + p] = cplx<T>{acc, im};
}

// ------------------------------------------------------------------ host-side helpers
template <typename F>
hipError_t dispatch_ns(int ns, F&& f) {
  switch (ns) {
    case 2: return f(std::integral_constant<int, 2>{});
    case 4: return f(std::integral_constant<int, 4>{});
    case 6: return f(std::integral_constant<int, 6>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 10: return f(std::integral_constant<int, 10>{});
    case 12: return f(std::integral_constant<int, 12>{});
    case 14: return f(std::integral_constant<int, 14>{});
    case 16: return f(std::integral_constant<int, 16>{});
    default: return hipErrorInvalidValue;
  }
}

template <typename T>
SIArgs<T> make_si_args(const GridOpArgs& ga) {
  const PointSet& ps = *ga.ps;
  const Geometry& g = *ga.g;
  SIArgs<T> a;
  a.perm = ps.perm;
  a.crec = (std::is_same<T, float>::value && g.dim == 2 && ps.has_crec) ? ps.crec : nullptr;
  a.cmax = nullptr;
  a.subprob = ps.subprob;
  a.nsub = ps.nsub;
  for (int d = 0; d < 3; ++d) {
    a.coord[d] = ps.coord[d];
    a.nf[d] = g.nf[d];
    a.bs[d] = g.bs[d];
    a.nb[d] = g.nb[d];
    a.tile[d] = g.tile[d];
    a.dscale[d] = (T)(-(double)g.nf[d] / kTwoPi);
  }
  a.M = ps.M;
  a.tile_cells = g.tile_cells;
  a.grid_cells = g.grid_cells();
  const double log2e = 1.4426950408889634074;
  a.es.beta_log2e = std::is_same<T, float>::value ? (T)(ga.beta * log2e) : (T)ga.beta;
  a.es.beta = (T)ga.beta;
  a.es.c = (T)ga.c;
  a.poly.on = 0;
  if (std::is_same<T, float>::value && g.ns == kEsPolyNs) a.poly = es_poly_table(g.ns, ga.beta, ga.c);
  return a;
}

// LDS bytes per tile cell: the spread tile holds two 8-byte accumulators per cell
// (64-bit fixed point or FP64), the gather tile one complex value.
// Tiles up to the device's per-block LDS limit (160 KB on gfx950) run tiled; bins are
// sized for the smaller occupancy budget first, but the bin-count cap of the sort can
// force larger tiles (3D grids of 256^3 cells at ns = 7: 23 x 15 x 15 cells, 83 KB).
// NUFFTX_LDS_TILE_MAX lowers the limit (read per call), which sends larger tiles to
// the global-memory kernels.
inline bool use_tiles(const Geometry& g, int prec, size_t& lds, bool spread = false) {
  const size_t cell = spread ? 16 : (prec == kF32 ? 8 : 16);
  lds = (size_t)g.tile_cells * cell;
  size_t cap = (size_t)lds_device_max_bytes();
  if (const char* e = std::getenv("NUFFTX_LDS_TILE_MAX")) cap = std::min(cap, (size_t)std::max(4096, std::atoi(e)));
  return lds <= cap;
}

}  // namespace nufftx
