// Stable bin-sort of nonuniform points (no reference counterpart: nufftax
// scatters unsorted, nufftax/core/spread.py:330-334).
//
// Contract: perm == argsort(key, stable) and bin_start == [0, cumsum(bincount(key))],
// where key = sum_d (floor(xs_d) div bs_d) * prod_{e<d} nb_e and xs is the FP64 fold.
//
// Counting sort, all integer work, in three launches (five above 1024 bins, see B):
//   A  k_keys: one block per chunk of points folds each point once, stores its
//      key and counts the keys in LDS (per-lane integer atomics); the block writes
//      its row of the histogram H[chunk][bin] (coalesced) and adds the row into
//      the sums of its segment of 16 rows, S[seg][bin] (global integer atomics).
//   B  k_seg_scan (one block): S -> exclusive prefixes over the segments (P) and
//      bin totals, the totals -> bin_start and, per bin, the first subproblem
//      (sub_start); S is cleared for the next sort.  Above 1024 bins the prefixes run
//      one thread per bin (k_seg_prefix), the bin scan in one block (k_bin_scan), and
//      k_row_offsets turns every histogram row into first output slots.
//   C  k_scatter: each block first fills its share of the subproblem descriptors
//      {bin, first sorted slot, count} (bins split into <= maxsub points) used by
//      the tiled spread/interp kernels; then the W waves of a chunk count their
//      share of the chunk's keys, turn the counts into per-wave bases (bin start +
//      segment prefix + rows of the segment before the chunk, or the row offsets of
//      k_row_offsets), and walk their keys
//      in index order, ranking equal keys inside each 64-key step (one comparison
//      when the step's keys agree, else one ballot per key bit), and write
//      perm[slot] = index.  Above 1024 bins without point records, k_scatter16 keeps
//      16-bit wave counters (two per LDS word) so that twice as many waves fit.
#include <algorithm>
#include <cstdlib>

#include "internal.hpp"

namespace nufftx {

namespace {

struct KeyArgs {
  const void* coord[3];
  int nf[3];
  int nb[3];
  double inv_bs[3];  // 1 / bs, rounded
};

template <typename CT, int D>
__device__ __forceinline__ void load_coords(const KeyArgs& ka, int64_t i, double (&x)[D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) x[d] = (double)static_cast<const CT*>(ka.coord[d])[i];
}

template <int D>
__device__ __forceinline__ int key_of(const KeyArgs& ka, const double (&x)[D]) {
  int key = 0, stride = 1;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const Folded f = fold_coord(x[d], ka.nf[d]);
    // cell div bs, exactly: (cell + 1/2) / bs is at least 1/(2 bs) away from an
    // integer, far more than the rounding error of the product
    key += (int)(((double)f.cell + 0.5) * ka.inv_bs[d]) * stride;
    stride *= ka.nb[d];
  }
  return key;
}

// Lanes of the calling wave whose key equals this lane's key (all lanes active).
// Valid lanes form a prefix of the wave.  When they all hold one key this is a
// single comparison; otherwise one ballot per key bit, O(log nbins).
__device__ __forceinline__ uint64_t match_key(int key, bool valid, int key_bits) {
  const uint64_t vmask = __ballot(valid);
  const int k0 = __shfl(key, 0, 64);
  if (__ballot(valid && key != k0) == 0) return vmask;
  uint64_t m = vmask;
  for (int b = 0; b < key_bits; ++b) {
    const bool bit = (key >> b) & 1;
    const uint64_t mb = __ballot(bit);
    m &= bit ? mb : ~mb;
  }
  return m;
}

constexpr int kStep = 64;          // points per wave step
constexpr int kGroup = 8;          // steps whose coordinates are loaded together

// ------------------------------------------------------------------ scan helpers
constexpr int kRowSeg = 16;  // histogram rows per scan segment

// Block-wide exclusive scan of one value per thread (blockDim.x a multiple of 64,
// at most 1024); returns the exclusive prefix and sets the block total.
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[16];
  const int lane = lane_id();
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int wbase = 0, tot = 0;
  for (int k = 0; k < nw; ++k) {
    const int s = wsum[k];
    if (k < w) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + incl - v;
}

struct ScanArgs {
  int32_t* H;          // [NB][nbins] chunk histograms
  int32_t* S;          // [nseg][nbins] segment sums (zero between sorts)
  int32_t* P;          // [nseg][nbins] exclusive prefixes over the segments
  int32_t* bin_start;  // [nbins + 1]
  int32_t* sub_start;  // [nbins + 1]
  int32_t* nsub;
  int4* subprob;
  int NB, nseg, nbins, M, maxsub, max_subprob;
  int rows_are_offsets;  // H already holds first output slots (k_row_offsets ran)
};

// S -> P for bin b (exclusive over the segments, S cleared); the bin total into bin_start[b]
__device__ __forceinline__ void seg_prefix_bin(const ScanArgs& sa, int b) {
  constexpr int kG = 16;  // segment sums loaded together (independent loads in flight)
  int run = 0;
  for (int g0 = 0; g0 < sa.nseg; g0 += kG) {
    int v[kG];
#pragma unroll
    for (int u = 0; u < kG; ++u) v[u] = g0 + u < sa.nseg ? sa.S[(int64_t)(g0 + u) * sa.nbins + b] : 0;
#pragma unroll
    for (int u = 0; u < kG; ++u) {
      if (g0 + u < sa.nseg) {
        const int64_t i = (int64_t)(g0 + u) * sa.nbins + b;
        sa.S[i] = 0;
        sa.P[i] = run;
        run += v[u];
      }
    }
  }
  sa.bin_start[b] = run;
}

// one block: bin totals (in bin_start) -> exclusive scan = bin_start; subproblems per bin ->
// sub_start; bin_start[nbins] = M; nsub.  Each thread takes a run of consecutive bins (16 at
// 16384 bins and 1024 threads): its sums, one block scan, then its run written out.
__device__ __forceinline__ void bin_scan(const ScanArgs& sa) {
  const int per = (sa.nbins + (int)blockDim.x - 1) / (int)blockDim.x;
  const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, sa.nbins);
  int cnt = 0, nsb = 0;
#pragma unroll 4
  for (int b = b0; b < b1; ++b) {
    const int c = sa.bin_start[b];
    cnt += c;
    nsb += (c + sa.maxsub - 1) / sa.maxsub;
  }
  int total, total_sub;
  int ex = block_excl_scan(cnt, total);
  int ex_sub = block_excl_scan(nsb, total_sub);
#pragma unroll 4
  for (int b = b0; b < b1; ++b) {
    const int c = sa.bin_start[b];  // this thread's own entries: read before it overwrites them
    sa.bin_start[b] = ex;
    sa.sub_start[b] = ex_sub;
    ex += c;
    ex_sub += (c + sa.maxsub - 1) / sa.maxsub;
  }
  if (threadIdx.x == 0) {
    sa.bin_start[sa.nbins] = sa.M;
    sa.sub_start[sa.nbins] = total_sub;
    *sa.nsub = min(total_sub, sa.max_subprob);
  }
}

// few bins: the segment prefixes and the bin scan in one block
__global__ void __launch_bounds__(1024) k_seg_scan(ScanArgs sa) {
  for (int b = threadIdx.x; b < sa.nbins; b += blockDim.x) seg_prefix_bin(sa, b);
  __syncthreads();
  bin_scan(sa);
}

// many bins: segment prefixes one thread per bin, then the bin scan in one block, then the
// chunk rows of the histogram turned into first output slots, H[r][b] -> slot of (bin b,
// chunk r), so that a k_scatter block reads one offset per bin
__global__ void __launch_bounds__(256) k_seg_prefix(ScanArgs sa) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < sa.nbins) seg_prefix_bin(sa, b);
}

__global__ void __launch_bounds__(1024) k_bin_scan(ScanArgs sa) { bin_scan(sa); }

__global__ void __launch_bounds__(256) k_row_offsets(ScanArgs sa) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  const int seg = blockIdx.y;
  if (b >= sa.nbins) return;
  const int r0 = seg * kRowSeg, r1 = min(sa.NB, r0 + kRowSeg);
  int run = sa.bin_start[b] + sa.P[(int64_t)seg * sa.nbins + b];
  for (int r = r0; r < r1; ++r) {
    const int t = sa.H[(int64_t)r * sa.nbins + b];
    sa.H[(int64_t)r * sa.nbins + b] = run;
    run += t;
  }
}

// k_keys: a block of 1024 threads folds the points of chunk blockIdx.x
// ([lo, lo + CH)), stores their keys and counts them.
template <typename CT, int D>
__global__ void __launch_bounds__(1024) k_keys(KeyArgs ka, int64_t M, int CH, int nbins, int32_t* __restrict__ keys,
                                               ScanArgs sa) {
  int32_t* __restrict__ H = sa.H;
  extern __shared__ int32_t cnt[];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * CH;
  const int64_t hi = min(M, lo + (int64_t)CH);
  const int stride = blockDim.x;
  for (int64_t base = lo; base < hi; base += (int64_t)stride * kGroup) {
    double x[kGroup][D];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + (int64_t)u * stride + threadIdx.x;
      if (i < hi) load_coords<CT, D>(ka, i, x[u]);
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + (int64_t)u * stride + threadIdx.x;
      if (i < hi) {
        const int key = key_of<D>(ka, x[u]);
        keys[i] = key;
        atomicAdd(&cnt[key], 1);
      }
    }
  }
  __syncthreads();
  int32_t* row = H + (int64_t)blockIdx.x * nbins;
  int32_t* seg = sa.S + (int64_t)(blockIdx.x / kRowSeg) * nbins;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    const int v = cnt[b];
    row[b] = v;
    if (v) atomicAdd(&seg[b], v);
  }
}

// Point records of 2D FP32 point sets, written at each point's sorted position so
// that the spread/interp kernels read them contiguously instead of gathering the
// coordinates through perm (a random gather costs a 128-byte line per value).
// Compact records (the LDS-tiled kernels), 8 bytes: one word per dimension,
//   word = o << 25 | u,  o = i0 - (bin origin - ns/2) in [0, bs] (support start in
//   the bin's tile, bs <= 127), u = round((t + ns/2 + 1) 2^24) with t = i0 - xs in
//   [-ns/2 - 1, -ns/2 + 1) from the FP64 fold (i0 = ceil(xs - ns/2));
// the point's index is perm at the same slot.  Full records (the matrix-core
// kernels), 16 bytes: {t_x, t_y, meta, i} with t in FP32, meta = ox | oy << 12 | s << 24
// and s the 4 x 4 sub-bin of the matrix-core geometry (bins of 4 x 4 sub-bins of
// sb = 16 - ns cells).
struct RecArgs {
  const void* coord[2];
  int nf[2], bs[2], nb0, ns, sb;
  float4* recs;  // full records, or null
  uint2* crec;   // compact records, or null
  int32_t* slot;  // inverse permutation (slot[i] = sorted position of point i), or null
};

__device__ __forceinline__ float4 point_record(const RecArgs& ra, int key, double x, double y, int i) {
  const int bx = key % ra.nb0, by = key / ra.nb0;
  const int bin_x0 = bx * ra.bs[0], bin_y0 = by * ra.bs[1];
  const int half = ra.ns / 2;
  const Folded fx = fold_coord(x, ra.nf[0]);
  const Folded fy = fold_coord(y, ra.nf[1]);
  const int i0x = support_start(fx, half), i0y = support_start(fy, half);
  const int ox = i0x - bin_x0 + half, oy = i0y - bin_y0 + half;  // in [0, bs]
  const int sub = ra.sb > 0 ? ((fy.cell - bin_y0) / ra.sb) * 4 + (fx.cell - bin_x0) / ra.sb : 0;
  return make_float4((float)((double)i0x - fx.xs), (float)((double)i0y - fy.xs),
                     __int_as_float(ox | (oy << 12) | (sub << 24)), __int_as_float(i));
}

__device__ __forceinline__ unsigned compact_word(double x, int nf, int bin0, int half) {
  const Folded f = fold_coord(x, nf);
  const int i0 = support_start(f, half);
  const double u = ((double)i0 - f.xs + (double)(half + 1)) * 16777216.0;  // 2^24
  const unsigned q = (unsigned)min(max(rint(u), 0.0), 33554431.0);      // 2^25 - 1
  return ((unsigned)(i0 - bin0 + half) << 25) | q;
}

__device__ __forceinline__ uint2 compact_record(const RecArgs& ra, int key, double x, double y) {
  const int bx = key % ra.nb0, by = key / ra.nb0;
  const int half = ra.ns / 2;
  return make_uint2(compact_word(x, ra.nf[0], bx * ra.bs[0], half), compact_word(y, ra.nf[1], by * ra.bs[1], half));
}

// k_scatter: W waves per chunk; wave w owns keys [lo + w*SC, lo + (w+1)*SC).
// REC: also write the 2D point records (coordinates read in index order), 1 = full,
// 2 = compact.
template <typename CT, int REC>
__global__ void __launch_bounds__(1024) k_scatter(const int32_t* __restrict__ keys, int64_t M, int SC, int nbins,
                                                  int key_bits, ScanArgs sa, int32_t* __restrict__ perm, RecArgs ra) {
  extern __shared__ int32_t wcnt[];  // [W][nbins]
  // subproblem descriptors; slot's bin is the last b with sub_start[b] <= slot
  {
    const int n = min(*sa.nsub, sa.max_subprob);
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < n; slot += gridDim.x * blockDim.x) {
      int lo = 0, hi = nbins;  // invariant: sub_start[lo] <= slot < sub_start[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sa.sub_start[mid] <= slot)
          lo = mid;
        else
          hi = mid;
      }
      const int k = slot - sa.sub_start[lo];
      const int st = sa.bin_start[lo], cnt = sa.bin_start[lo + 1] - st;
      sa.subprob[slot] = make_int4(lo, st + k * sa.maxsub, min(sa.maxsub, cnt - k * sa.maxsub), 0);
    }
  }
  const int W = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int32_t* my = wcnt + wave * nbins;
  for (int b = lane; b < nbins; b += 64) my[b] = 0;
  const int64_t lo = (int64_t)blockIdx.x * W * SC + (int64_t)wave * SC;
  const int64_t hi = min(M, lo + (int64_t)SC);
  // count this wave's keys (LDS integer atomics on the wave's own counters)
  for (int64_t base = lo; base < hi; base += kStep * kGroup) {
    int k[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      k[u] = i < hi ? keys[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u)
      if (k[u] >= 0) atomicAdd(&my[k[u]], 1);
  }
  __syncthreads();
  // per-wave bases: first slot of (bin, chunk) = bin start + prefix of the earlier
  // segments + the rows of this segment before the chunk; then earlier waves' counts
  const int chunk = (int)blockIdx.x, seg = chunk / kRowSeg;
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) {
    int run;
    if (sa.rows_are_offsets) {
      run = sa.H[(int64_t)chunk * nbins + b];
    } else {
      run = sa.bin_start[b] + sa.P[(int64_t)seg * nbins + b];
      for (int r = seg * kRowSeg; r < chunk; ++r) run += sa.H[(int64_t)r * nbins + b];
    }
    for (int w = 0; w < W; ++w) {
      const int t = wcnt[w * nbins + b];
      wcnt[w * nbins + b] = run;
      run += t;
    }
  }
  __syncthreads();
  // rank in index order and write the permutation
  for (int64_t base = lo; base < hi; base += kStep * kGroup) {
    int k[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      k[u] = i < hi ? keys[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      const bool valid = i < hi;
      const int key = k[u];
      const uint64_t m = match_key(key, valid, key_bits);
      const int rank = popc_below(m);
      const int b0 = my[key];
      if (valid && rank == 0) my[key] = b0 + (int)__popcll(m);
      const int pos = b0 + rank;
      if (valid && pos >= 0 && (int64_t)pos < M) {
        perm[pos] = (int32_t)i;
        if (ra.slot) ra.slot[i] = pos;
        if constexpr (REC != 0) {
          const double x = (double)static_cast<const CT*>(ra.coord[0])[i];
          const double y = (double)static_cast<const CT*>(ra.coord[1])[i];
          if constexpr (REC == 1)
            ra.recs[pos] = point_record(ra, key, x, y, (int)i);
          else
            ra.crec[pos] = compact_record(ra, key, x, y);
        }
      }
    }
  }
}

// k_scatter16: k_scatter without records for more than 1024 bins, with 16-bit wave
// counters (two per LDS word), so that twice as many waves fit a chunk (16384 bins: 4 waves
// instead of 2).  A wave's counters hold offsets inside the chunk (below 65536: the host
// keeps chunks that small); the first slot of (chunk, bin) comes from the row offsets.
// Counting adds 1 << 16 (key & 1) to word key / 2 with a 32-bit LDS atomic; ranking lets the
// lowest lane of each key add its count the same way and hands the old offset to the others.
template <typename CT>
__global__ void __launch_bounds__(1024) k_scatter16(const int32_t* __restrict__ keys, int64_t M, int SC, int nbins,
                                                    int key_bits, ScanArgs sa, int32_t* __restrict__ perm,
                                                    int32_t* __restrict__ slot) {
  extern __shared__ uint32_t wc2[];  // [W][(nbins + 1) / 2]
  {
    const int n = min(*sa.nsub, sa.max_subprob);
    for (int slot = blockIdx.x * blockDim.x + threadIdx.x; slot < n; slot += gridDim.x * blockDim.x) {
      int lo = 0, hi = nbins;  // invariant: sub_start[lo] <= slot < sub_start[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (sa.sub_start[mid] <= slot)
          lo = mid;
        else
          hi = mid;
      }
      const int k = slot - sa.sub_start[lo];
      const int st = sa.bin_start[lo], cnt = sa.bin_start[lo + 1] - st;
      sa.subprob[slot] = make_int4(lo, st + k * sa.maxsub, min(sa.maxsub, cnt - k * sa.maxsub), 0);
    }
  }
  const int W = blockDim.x >> 6;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int nw2 = (nbins + 1) >> 1;
  uint32_t* my = wc2 + (int64_t)wave * nw2;
  for (int b = lane; b < nw2; b += 64) my[b] = 0u;
  const int64_t lo = (int64_t)blockIdx.x * W * SC + (int64_t)wave * SC;
  const int64_t hi = min(M, lo + (int64_t)SC);
  for (int64_t base = lo; base < hi; base += kStep * kGroup) {
    int k[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      k[u] = i < hi ? keys[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u)
      if (k[u] >= 0) atomicAdd(&my[k[u] >> 1], 1u << (16 * (k[u] & 1)));
  }
  __syncthreads();
  // offsets inside the chunk: earlier waves' counts of the bin (both halves at once: no half
  // reaches 65536, so no carry crosses)
  for (int b2 = threadIdx.x; b2 < nw2; b2 += blockDim.x) {
    uint32_t run = 0u;
    for (int w = 0; w < W; ++w) {
      const uint32_t t = wc2[(int64_t)w * nw2 + b2];
      wc2[(int64_t)w * nw2 + b2] = run;
      run += t;
    }
  }
  __syncthreads();
  const int32_t* __restrict__ first = sa.H + (int64_t)blockIdx.x * nbins;  // k_row_offsets
  for (int64_t base = lo; base < hi; base += kStep * kGroup) {
    int k[kGroup];
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      k[u] = i < hi ? keys[i] : 0;
    }
#pragma unroll
    for (int u = 0; u < kGroup; ++u) {
      const int64_t i = base + u * kStep + lane;
      const bool valid = i < hi;
      const int key = k[u];
      const uint64_t m = match_key(key, valid, key_bits);
      const int rank = popc_below(m);
      const int sh = 16 * (key & 1);
      uint32_t old = 0u;
      if (valid && rank == 0) old = atomicAdd(&my[key >> 1], (uint32_t)__popcll(m) << sh);
      const int leader = m ? __ffsll((long long)m) - 1 : lane;
      const int off = __shfl((int)((old >> sh) & 0xFFFFu), leader, 64);
      const int pos = first[key] + off + rank;
      if (valid && pos >= 0 && (int64_t)pos < M) {
        perm[pos] = (int32_t)i;
        if (slot) slot[i] = pos;
      }
    }
  }
}

// Compact records of a 2D point set through the inverse permutation: crec[slot[i]] = the
// record of point i (coordinates, keys and slots read in index order, 8-byte records
// scattered).  A pass of its own instead of a write inside the sort's ranking loop: the
// stream of independent scattered stores keeps more writes in flight (measured on the 2D
// radial workload, 4M points: the scatter with records took 63 us, the plain scatter and
// this pass take 27 + 25 us).
template <typename CT>
__global__ void __launch_bounds__(256) k_pack_crec(RecArgs ra, const int32_t* __restrict__ keys, int64_t M) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) {
    const int pos = ra.slot[i];
    if ((unsigned)pos >= (unsigned)M) continue;
    const double x = (double)static_cast<const CT*>(ra.coord[0])[i];
    const double y = (double)static_cast<const CT*>(ra.coord[1])[i];
    ra.crec[pos] = compact_record(ra, keys[i], x, y);
  }
}

// Packed records of a 2D FP32 point set for the matrix-core window spread: point i of
// transform t goes to packed[t][slot[i]] = {x word, y word, re c, im c} (compact words as
// in compact_record, tile-relative to the point's bin from its stored key).  Coordinates,
// keys, slots and strengths are read in index order; the 16-byte records are scattered.
template <typename CT>
__global__ void __launch_bounds__(256) k_pack2d(RecArgs ra, const int32_t* __restrict__ keys, int64_t M,
                                                int64_t n_trans, const float2* __restrict__ c,
                                                uint4* __restrict__ packed, uint4* __restrict__ zero, int64_t nzero) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  // the output grid of the spread that follows, cleared here instead of by a launch of its own
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nzero; i += stride)
    zero[i] = make_uint4(0u, 0u, 0u, 0u);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) {
    const int pos = ra.slot[i];
    if ((unsigned)pos >= (unsigned)M) continue;
    const double x = (double)static_cast<const CT*>(ra.coord[0])[i];
    const double y = (double)static_cast<const CT*>(ra.coord[1])[i];
    const uint2 w = compact_record(ra, keys[i], x, y);
    const bool fin = isfinite(x) && isfinite(y);
    for (int64_t t = 0; t < n_trans; ++t) {
      float2 v = c[t * M + i];
      if (!fin) v = make_float2(NAN, NAN);
      packed[t * M + pos] = make_uint4(w.x, w.y, __float_as_uint(v.x), __float_as_uint(v.y));
    }
  }
}

// Packed words of a 3D point set for the 3D window spread: packed[slot[i]] = {x, y, z word}
// (the layout of compact_word, tile-relative to the point's bin from its stored key) and
// cs[t][slot[i]] = c[t][i], NaN for a non-finite coordinate.  Index-order reads, scattered
// writes; the first loop clears the spread's output grid when asked.
struct Pack3Args {
  const void* coord[3];
  int nf[3], bs[3], nb[3], ns;
  const int32_t* slot;
};

template <typename CT>
__global__ void __launch_bounds__(256) k_pack3d(Pack3Args pa, const int32_t* __restrict__ keys, int64_t M,
                                                int64_t n_trans, const float2* __restrict__ c,
                                                uint4* __restrict__ packed, float2* __restrict__ cs,
                                                uint4* __restrict__ zero, int64_t nzero) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nzero; i += stride)
    zero[i] = make_uint4(0u, 0u, 0u, 0u);
  const int half = pa.ns / 2;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < M; i += stride) {
    const int pos = pa.slot[i];
    if ((unsigned)pos >= (unsigned)M) continue;
    int key = keys[i];
    unsigned w[3];
    bool fin = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const int b = key % pa.nb[d];
      key /= pa.nb[d];
      const double x = (double)static_cast<const CT*>(pa.coord[d])[i];
      fin = fin && isfinite(x);
      w[d] = compact_word(x, pa.nf[d], b * pa.bs[d], half);
    }
    packed[pos] = make_uint4(w[0], w[1], w[2], 0u);
    for (int64_t t = 0; t < n_trans; ++t) {
      float2 v = c[t * M + i];
      if (!fin) v = make_float2(NAN, NAN);
      cs[t * M + pos] = v;
    }
  }
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_empty_sort(int nbins, int32_t* bin_start, int32_t* nsub) {
  for (int b = threadIdx.x; b <= nbins; b += blockDim.x) bin_start[b] = 0;
  if (threadIdx.x == 0) *nsub = 0;
}

template <typename CT, int D>
void launch_sort_passes(const KeyArgs& ka, PointSet& ps, int nbins, int W, int SC, bool c16, const RecArgs& ra,
                        hipStream_t s) {
  int key_bits = 0;
  while ((1 << key_bits) < nbins) ++key_bits;
  const int CH = W * SC;
  ScanArgs sa;
  sa.H = ps.hist;
  sa.S = ps.scan_sum;
  sa.P = ps.scan_part;
  sa.bin_start = ps.bin_start;
  sa.sub_start = ps.sub_start;
  sa.nsub = ps.nsub;
  sa.subprob = ps.subprob;
  sa.NB = ps.NB;
  sa.nseg = (ps.NB + kRowSeg - 1) / kRowSeg;
  sa.nbins = nbins;
  sa.M = (int)ps.M;
  sa.maxsub = ps.maxsub;
  sa.max_subprob = ps.max_subprob;
  const size_t lds_keys = (size_t)nbins * sizeof(int32_t);
  // more than one bin per scatter thread: offsets of every (chunk, bin) in a pass of their own
  sa.rows_are_offsets = nbins > 1024 ? 1 : 0;
  hipLaunchKernelGGL((k_keys<CT, D>), dim3(ps.NB), dim3(1024), lds_keys, s, ka, ps.M, CH, nbins, ps.keys, sa);
  if (sa.rows_are_offsets) {
    hipLaunchKernelGGL(k_seg_prefix, dim3((nbins + 255) / 256), dim3(256), 0, s, sa);
    hipLaunchKernelGGL(k_bin_scan, dim3(1), dim3(1024), 0, s, sa);
    hipLaunchKernelGGL(k_row_offsets, dim3((nbins + 255) / 256, sa.nseg), dim3(256), 0, s, sa);
  } else {
    hipLaunchKernelGGL(k_seg_scan, dim3(1), dim3(1024), 0, s, sa);
  }
  if (c16) {
    const size_t l16 = (size_t)W * (size_t)((nbins + 1) / 2) * sizeof(uint32_t);
    if (l16 > 65536)
      (void)hipFuncSetAttribute((const void*)k_scatter16<CT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l16);
    hipLaunchKernelGGL((k_scatter16<CT>), dim3(ps.NB), dim3(64 * W), l16, s, ps.keys, ps.M, SC, nbins, key_bits, sa,
                       ps.perm, ra.slot);
    return;
  }
  const size_t lds = (size_t)W * nbins * sizeof(int32_t);
  if (lds > 65536) {
    (void)hipFuncSetAttribute((const void*)k_scatter<CT, 0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)k_scatter<CT, 1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    (void)hipFuncSetAttribute((const void*)k_scatter<CT, 2>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  }
  if (D == 2 && ra.recs)
    hipLaunchKernelGGL((k_scatter<CT, 1>), dim3(ps.NB), dim3(64 * W), lds, s, ps.keys, ps.M, SC, nbins, key_bits, sa,
                       ps.perm, ra);
  else if (D == 2 && ra.crec && !ra.slot)
    hipLaunchKernelGGL((k_scatter<CT, 2>), dim3(ps.NB), dim3(64 * W), lds, s, ps.keys, ps.M, SC, nbins, key_bits, sa,
                       ps.perm, ra);
  else if (D == 2 && ra.crec) {
    hipLaunchKernelGGL((k_scatter<CT, 0>), dim3(ps.NB), dim3(64 * W), lds, s, ps.keys, ps.M, SC, nbins, key_bits, sa,
                       ps.perm, ra);
    const unsigned nb = (unsigned)std::min<int64_t>((ps.M + 255) / 256, 8192);
    hipLaunchKernelGGL((k_pack_crec<CT>), dim3(nb), dim3(256), 0, s, ra, ps.keys, ps.M);
  } else
    hipLaunchKernelGGL((k_scatter<CT, 0>), dim3(ps.NB), dim3(64 * W), lds, s, ps.keys, ps.M, SC, nbins, key_bits, sa,
                       ps.perm, ra);
}

}  // namespace

SortSizes sort_sizes(int64_t M, const Geometry& g, const SortTuning& tune, int rec_kind) {
  SortSizes z;
  const int nbins = g.nbins;
  // waves per chunk (up to 16): their LDS counters (W * nbins ints) must fit the device's LDS
  const int64_t lds_max = lds_device_max_bytes();
  int W = 16;
  while (W > 1 && (int64_t)W * nbins * (int64_t)sizeof(int32_t) > lds_max) W /= 2;
  // points per wave: chunk / W, and enough that the histogram fits hist_cap
  const int64_t need = (M * (int64_t)nbins + tune.hist_cap - 1) / tune.hist_cap;  // min points per chunk
  auto wave_points = [&](int w) {
    int64_t sc = std::max<int64_t>(tune.chunk / w, 64);
    sc = std::max<int64_t>(sc, (need + w - 1) / w);
    sc = (sc + 63) / 64 * 64;
    return std::min<int64_t>(sc, 1 << 24);
  };
  int64_t SC = wave_points(W);
  // 16-bit wave counters when 32-bit ones leave fewer than 16 waves per chunk (more than 1024
  // bins, no records, chunks below 65536 points)
  bool c16 = false;
  const bool want_rec = (rec_kind == 1 || rec_kind == 2) && g.dim == 2;
  if (W < 16 && nbins > 1024 && !want_rec && !std::getenv("NUFFTX_SORT_C32")) {
    int W16 = 16;
    while (W16 > 1 && (int64_t)W16 * ((nbins + 1) / 2) * (int64_t)sizeof(uint32_t) > lds_max) W16 /= 2;
    const int64_t sc16 = wave_points(W16);
    if (W16 > W && (int64_t)W16 * sc16 <= 65535) {
      W = W16;
      SC = sc16;
      c16 = true;
    }
  }
  z.W = W;
  z.SC = SC;
  z.c16 = c16;
  z.NB = std::max<int64_t>(1, (M + W * SC - 1) / (W * SC));
  z.nhist = (int64_t)nbins * z.NB;
  z.nparts = ((z.NB + kRowSeg - 1) / kRowSeg) * nbins;
  z.max_sub = nbins + (M + tune.maxsub - 1) / tune.maxsub;
  return z;
}

hipError_t pointset_sort(PointSet& ps, const Geometry& g, const SortTuning& tune, hipStream_t s, std::string& err,
                         int rec_kind, int sub_sb, bool want_slot, int64_t n_trans) {
  hipError_t e;
  const int nbins = g.nbins;
  if ((int64_t)nbins * sizeof(int32_t) > 65536) {
    err = "bin count exceeds the LDS histogram capacity (16384 bins)";
    return hipErrorInvalidValue;
  }
  if (ps.M > (int64_t)INT32_MAX - 1) {
    err = "number of points exceeds 2^31 - 2";
    return hipErrorInvalidValue;
  }
  ps.maxsub = tune.maxsub;
  const int64_t M = ps.M;
  const SortSizes z = sort_sizes(M, g, tune, rec_kind);
  const int W = z.W;
  const int64_t SC = z.SC;
  const bool c16 = z.c16;
  const int64_t NB = z.NB;
  ps.NB = (int)NB;
  ps.CH = (int)(W * SC);
  const int64_t nhist = z.nhist;
  const int64_t nparts = z.nparts;
  const int64_t max_sub = z.max_sub;
  if ((e = grow_buffer(ps.perm, ps.cap_M, M, ps.retired)) != hipSuccess) return e;
  const bool recs = rec_kind == 1 && g.dim == 2 && M > 0;
  const bool crec = rec_kind == 2 && g.dim == 2 && M > 0 && g.bs[0] <= 127 && g.bs[1] <= 127;
  if (recs && (e = grow_buffer(ps.recs, ps.cap_recs, M, ps.retired)) != hipSuccess) return e;
  if (crec && (e = grow_buffer(ps.crec, ps.cap_crec, M, ps.retired)) != hipSuccess) return e;
  // the inverse permutation also serves the compact records (written by a pass of their own)
  const bool slot2 = (want_slot || crec) && g.dim == 2 && M > 0 && g.bs[0] <= 127 && g.bs[1] <= 127;
  const bool slot3 = want_slot && g.dim == 3 && M > 0 && g.tile[0] <= 127 && g.tile[1] <= 127 && g.tile[2] <= 127;
  const bool slot = slot2 || slot3;
  if (slot && (e = grow_buffer(ps.slot, ps.cap_slot, M, ps.retired)) != hipSuccess) return e;
  if (slot2 && n_trans > 0 &&
      (e = grow_buffer(ps.packed, ps.cap_packed, M * n_trans, ps.retired)) != hipSuccess)
    return e;
  if (slot3 && n_trans > 0 &&
      ((e = grow_buffer(ps.packed, ps.cap_packed, M, ps.retired)) != hipSuccess ||
       (e = grow_buffer(ps.cs, ps.cap_cs, M * n_trans, ps.retired)) != hipSuccess))
    return e;
  ps.has_recs = false;
  ps.has_crec = false;
  ps.has_slot = false;
  ps.recs_subbin = false;
  if ((e = grow_buffer(ps.keys, ps.cap_keys, M, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.sub_start, ps.cap_sub_start, (int64_t)nbins + 1, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.hist, ps.cap_hist, nhist, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.scan_part, ps.cap_part, nparts, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.bin_start, ps.cap_bins, (int64_t)nbins + 1, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.subprob, ps.cap_sub, max_sub, ps.retired)) != hipSuccess) return e;
  if (!ps.nsub && (e = dev_malloc((void**)&ps.nsub, sizeof(int32_t))) != hipSuccess) return e;
  if (nparts > ps.cap_sum) {
    // segment sums start at zero; k_seg_scan leaves them at zero
    if ((e = grow_buffer(ps.scan_sum, ps.cap_sum, nparts, ps.retired)) != hipSuccess) return e;
    const unsigned nz = (unsigned)std::min<int64_t>((ps.cap_sum + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero_i32, dim3(nz), dim3(256), 0, s, ps.scan_sum, ps.cap_sum);
  }
  ps.max_subprob = (int)max_sub;

  if (M == 0) {
    hipLaunchKernelGGL(k_empty_sort, dim3(1), dim3(256), 0, s, nbins, ps.bin_start, ps.nsub);
    return hipGetLastError();
  }
  KeyArgs ka;
  for (int d = 0; d < 3; ++d) {
    ka.coord[d] = ps.coord[d];
    ka.nf[d] = g.nf[d];
    ka.nb[d] = g.nb[d];
    ka.inv_bs[d] = 1.0 / (double)g.bs[d];
  }
  const int isc = (int)SC;
  RecArgs ra;
  ra.coord[0] = ps.coord[0];
  ra.coord[1] = ps.coord[1];
  ra.nf[0] = g.nf[0];
  ra.nf[1] = g.nf[1];
  ra.bs[0] = g.bs[0];
  ra.bs[1] = g.bs[1];
  ra.nb0 = g.nb[0];
  ra.ns = g.ns;
  ra.sb = sub_sb;
  ra.recs = recs ? ps.recs : nullptr;
  ra.crec = crec ? ps.crec : nullptr;
  ra.slot = slot ? ps.slot : nullptr;
  if (ps.coord_f64) {
    if (g.dim == 1) launch_sort_passes<double, 1>(ka, ps, nbins, W, isc, c16, ra, s);
    if (g.dim == 2) launch_sort_passes<double, 2>(ka, ps, nbins, W, isc, c16, ra, s);
    if (g.dim == 3) launch_sort_passes<double, 3>(ka, ps, nbins, W, isc, c16, ra, s);
  } else {
    if (g.dim == 1) launch_sort_passes<float, 1>(ka, ps, nbins, W, isc, c16, ra, s);
    if (g.dim == 2) launch_sort_passes<float, 2>(ka, ps, nbins, W, isc, c16, ra, s);
    if (g.dim == 3) launch_sort_passes<float, 3>(ka, ps, nbins, W, isc, c16, ra, s);
  }
  e = hipGetLastError();
  if (e == hipSuccess && recs) ps.has_recs = true;
  if (e == hipSuccess && crec) ps.has_crec = true;
  if (e == hipSuccess && slot) ps.has_slot = true;
  return e;
}

hipError_t pointset_pack2d(PointSet& ps, const Geometry& g, const void* c, int64_t n_trans, hipStream_t s,
                           void* zero, int64_t zero_bytes) {
  if (!ps.has_slot || g.dim != 2 || zero_bytes % 16 != 0) return hipErrorInvalidValue;
  if (ps.M == 0) return hipSuccess;
  // sized by setpts for the plans that spread; a point set that is only interpolated
  // (type 2) gets its buffer on its first spread (the reverse-mode derivative)
  hipError_t e = grow_buffer(ps.packed, ps.cap_packed, ps.M * n_trans, ps.retired);
  if (e != hipSuccess) return e;
  RecArgs ra{};
  ra.coord[0] = ps.coord[0];
  ra.coord[1] = ps.coord[1];
  ra.nf[0] = g.nf[0];
  ra.nf[1] = g.nf[1];
  ra.bs[0] = g.bs[0];
  ra.bs[1] = g.bs[1];
  ra.nb0 = g.nb[0];
  ra.ns = g.ns;
  ra.slot = ps.slot;
  const int64_t nzero = zero ? zero_bytes / 16 : 0;
  const unsigned nb = (unsigned)std::min<int64_t>((std::max<int64_t>(ps.M, nzero) + 255) / 256, 8192);
  if (ps.coord_f64)
    hipLaunchKernelGGL(k_pack2d<double>, dim3(nb), dim3(256), 0, s, ra, ps.keys, ps.M, n_trans,
                       static_cast<const float2*>(c), ps.packed, static_cast<uint4*>(zero), nzero);
  else
    hipLaunchKernelGGL(k_pack2d<float>, dim3(nb), dim3(256), 0, s, ra, ps.keys, ps.M, n_trans,
                       static_cast<const float2*>(c), ps.packed, static_cast<uint4*>(zero), nzero);
  return hipGetLastError();
}

hipError_t pointset_pack3d(PointSet& ps, const Geometry& g, const void* c, int64_t n_trans, hipStream_t s,
                           void* zero, int64_t zero_bytes) {
  if (!ps.has_slot || g.dim != 3 || zero_bytes % 16 != 0) return hipErrorInvalidValue;
  if (ps.M == 0) return hipSuccess;
  hipError_t e;
  if ((e = grow_buffer(ps.packed, ps.cap_packed, ps.M, ps.retired)) != hipSuccess) return e;
  if ((e = grow_buffer(ps.cs, ps.cap_cs, ps.M * n_trans, ps.retired)) != hipSuccess) return e;
  Pack3Args pa{};
  for (int d = 0; d < 3; ++d) {
    pa.coord[d] = ps.coord[d];
    pa.nf[d] = g.nf[d];
    pa.bs[d] = g.bs[d];
    pa.nb[d] = g.nb[d];
  }
  pa.ns = g.ns;
  pa.slot = ps.slot;
  const int64_t nzero = zero ? zero_bytes / 16 : 0;
  const unsigned nb = (unsigned)std::min<int64_t>((std::max<int64_t>(ps.M, nzero) + 255) / 256, 8192);
  if (ps.coord_f64)
    hipLaunchKernelGGL(k_pack3d<double>, dim3(nb), dim3(256), 0, s, pa, ps.keys, ps.M, n_trans,
                       static_cast<const float2*>(c), ps.packed, ps.cs, static_cast<uint4*>(zero), nzero);
  else
    hipLaunchKernelGGL(k_pack3d<float>, dim3(nb), dim3(256), 0, s, pa, ps.keys, ps.M, n_trans,
                       static_cast<const float2*>(c), ps.packed, ps.cs, static_cast<uint4*>(zero), nzero);
  return hipGetLastError();
}

void pointset_free(PointSet& ps) {
  void* bufs[] = {ps.perm, ps.keys, ps.hist, ps.scan_part, ps.scan_sum, ps.bin_start, ps.sub_start, ps.subprob,
                  ps.nsub, ps.recs, ps.crec, ps.sub_off, ps.cmax, ps.slot, ps.packed, ps.cs};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  for (void* p : ps.retired) (void)hipFree(p);
  ps.retired.clear();
  ps.perm = nullptr;
  ps.keys = nullptr;
  ps.recs = nullptr;
  ps.crec = nullptr;
  ps.has_crec = false;
  ps.cap_crec = 0;
  ps.sub_off = nullptr;
  ps.cmax = nullptr;
  ps.slot = nullptr;
  ps.packed = nullptr;
  ps.cs = nullptr;
  ps.has_slot = false;
  ps.cap_slot = ps.cap_packed = ps.cap_cs = 0;
  ps.has_recs = false;
  ps.recs_subbin = false;
  ps.cap_recs = ps.cap_sub_off = ps.cap_cmax = 0;
  ps.sub_start = nullptr;
  ps.hist = nullptr;
  ps.scan_part = nullptr;
  ps.scan_sum = nullptr;
  ps.cap_sum = 0;
  ps.bin_start = nullptr;
  ps.subprob = nullptr;
  ps.nsub = nullptr;
  ps.cap_M = ps.cap_keys = ps.cap_hist = ps.cap_part = ps.cap_bins = ps.cap_sub = ps.cap_sub_start = 0;
}

}  // namespace nufftx
