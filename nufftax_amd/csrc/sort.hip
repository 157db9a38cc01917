// Stable bin-sort of nonuniform points (no reference counterpart: nufftax
// scatters unsorted, nufftax/core/spread.py:330-334).
//
// Contract: perm == argsort(key, stable) and bin_start == [0, cumsum(bincount(key))],
// where key = sum_d floor(xs_d)/bs_d * prod_{e<d} nb_e and xs is the FP64 fold.
//
// Counting sort in three passes, all integer work:
//   A  per-chunk histograms (LDS counters)        -> hist[bin][chunk]
//   B  exclusive scan of hist in (bin, chunk) order -> first slot of every (bin, chunk)
//   C  one wave per chunk walks its points in index order, ranks equal keys
//      inside each 64-point step with ballots, and writes perm[slot] = index.
// A subproblem list (bins split into <= maxsub points) is then built for the
// LDS-tiled spread/interp kernels.
#include <algorithm>

#include "internal.hpp"

namespace nufftx {

namespace {

struct KeyArgs {
  const void* coord[3];
  int nf[3];
  int bs[3];
  int nb[3];
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
    Folded f = fold_coord(x[d], ka.nf[d]);
    key += (f.cell / ka.bs[d]) * stride;
    stride *= ka.nb[d];
  }
  return key;
}

// Lanes of the calling wave whose key equals this lane's key (all lanes active),
// from one ballot per key bit: O(log nbins), no memory traffic.
__device__ __forceinline__ uint64_t match_key(int key, bool valid, int key_bits) {
  uint64_t m = __ballot(valid);
  for (int b = 0; b < key_bits; ++b) {
    const bool bit = (key >> b) & 1;
    const uint64_t mb = __ballot(bit);
    m &= bit ? mb : ~mb;
  }
  return m;
}

constexpr int kHistThreads = 256;
constexpr int kHistUnroll = 8;

template <typename CT, int D>
__global__ void __launch_bounds__(kHistThreads) k_hist(KeyArgs ka, int64_t M, int CH, int NB, int nbins,
                                                       int key_bits, int32_t* __restrict__ hist) {
  extern __shared__ int32_t cnt[];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * CH;
  const int64_t hi = min(M, lo + (int64_t)CH);
  for (int64_t base = lo; base < hi; base += (int64_t)kHistThreads * kHistUnroll) {
    double x[kHistUnroll][D];
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      int64_t i = base + (int64_t)u * kHistThreads + threadIdx.x;
      if (i < hi) load_coords<CT, D>(ka, i, x[u]);
    }
#pragma unroll
    for (int u = 0; u < kHistUnroll; ++u) {
      const int64_t i = base + (int64_t)u * kHistThreads + threadIdx.x;
      const bool valid = i < hi;
      const int key = valid ? key_of<D>(ka, x[u]) : 0;
      const uint64_t m = match_key(key, valid, key_bits);
      if (valid && popc_below(m) == 0) atomicAdd(&cnt[key], (int)__popcll(m));
    }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[(int64_t)b * NB + blockIdx.x] = cnt[b];
}

constexpr int kScatUnroll = 8;

template <typename CT, int D>
__global__ void __launch_bounds__(64) k_scatter(KeyArgs ka, int64_t M, int CH, int NB, int nbins, int key_bits,
                                                const int32_t* __restrict__ offs, int32_t* __restrict__ perm) {
  extern __shared__ int32_t base[];
  const int lane = threadIdx.x;
  for (int b = lane; b < nbins; b += 64) base[b] = offs[(int64_t)b * NB + blockIdx.x];
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * CH;
  const int64_t hi = min(M, lo + (int64_t)CH);
  for (int64_t grp = lo; grp < hi; grp += 64 * kScatUnroll) {
    double x[kScatUnroll][D];
#pragma unroll
    for (int u = 0; u < kScatUnroll; ++u) {
      int64_t i = grp + u * 64 + lane;
      if (i < hi) load_coords<CT, D>(ka, i, x[u]);
    }
#pragma unroll
    for (int u = 0; u < kScatUnroll; ++u) {
      const int64_t i = grp + u * 64 + lane;
      const bool valid = i < hi;
      const int key = valid ? key_of<D>(ka, x[u]) : 0;
      const uint64_t m = match_key(key, valid, key_bits);
      const int rank = popc_below(m);
      const int b0 = base[key];
      if (valid && rank == 0) base[key] = b0 + (int)__popcll(m);
      const int pos = b0 + rank;
      if (valid && pos >= 0 && (int64_t)pos < M) perm[pos] = (int32_t)i;
    }
  }
}

// ------------------------------------------------------------------ scan
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Block-wide exclusive scan of one value per thread; returns the block total.
template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int& total) {
  __shared__ int wsum[NT / 64];
  const int lane = lane_id();
  const int w = threadIdx.x >> 6;
  int incl = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(incl, off, 64);
    if (lane >= off) incl += t;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int wbase = 0;
  int tot = 0;
#pragma unroll
  for (int k = 0; k < NT / 64; ++k) {
    int s = wsum[k];
    if (k < w) wbase += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return wbase + incl - v;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const int32_t* __restrict__ a, int64_t n,
                                                              int32_t* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) s += a[i];
  }
  int total;
  block_excl_scan<kScanThreads>(s, total);
  if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// single block: exclusive scan of the partials in place
__global__ void __launch_bounds__(1024) k_scan_partials(int32_t* __restrict__ part, int n) {
  int carry = 0;
  for (int base = 0; base < n; base += 1024) {
    int i = base + threadIdx.x;
    int v = i < n ? part[i] : 0;
    int total;
    int ex = block_excl_scan<1024>(v, total);
    if (i < n) part[i] = carry + ex;
    carry += total;
  }
}

__global__ void __launch_bounds__(kScanThreads) k_scan_final(int32_t* __restrict__ a, int64_t n,
                                                             const int32_t* __restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    v[k] = i < n ? a[i] : 0;
    s += v[k];
  }
  int total;
  int run = block_excl_scan<kScanThreads>(s, total) + part[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) a[i] = run;
    run += v[k];
  }
}

// single block: bin starts and subproblem descriptors {bin, first sorted slot, count}
__global__ void __launch_bounds__(1024) k_subprob(const int32_t* __restrict__ offs, int NB, int nbins, int M,
                                                  int maxsub, int max_subprob, int32_t* __restrict__ bin_start,
                                                  int4* __restrict__ subprob, int32_t* __restrict__ nsub) {
  int carry = 0;
  for (int base = 0; base < nbins; base += 1024) {
    const int b = base + threadIdx.x;
    int st = 0, cnt = 0;
    if (b < nbins) {
      st = offs[(int64_t)b * NB];
      const int en = (b + 1 < nbins) ? offs[(int64_t)(b + 1) * NB] : M;
      cnt = en - st;
      bin_start[b] = st;
    }
    const int nsb = (cnt + maxsub - 1) / maxsub;
    int total;
    const int ex = block_excl_scan<1024>(nsb, total) + carry;
    for (int k = 0; k < nsb; ++k) {
      const int slot = ex + k;
      if (slot < max_subprob) subprob[slot] = make_int4(b, st + k * maxsub, min(maxsub, cnt - k * maxsub), 0);
    }
    carry += total;
  }
  if (threadIdx.x == 0) {
    bin_start[nbins] = M;
    *nsub = min(carry, max_subprob);
  }
}

__global__ void k_empty_sort(int nbins, int32_t* bin_start, int32_t* nsub) {
  for (int b = threadIdx.x; b <= nbins; b += blockDim.x) bin_start[b] = 0;
  if (threadIdx.x == 0) *nsub = 0;
}

template <typename T>
hipError_t grow(T*& p, int64_t& cap, int64_t need) {
  if (need <= cap && p) return hipSuccess;
  if (p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return e;
    p = nullptr;
  }
  int64_t n = std::max<int64_t>(need, 1);
  hipError_t e = hipMalloc((void**)&p, (size_t)n * sizeof(T));
  if (e != hipSuccess) {
    cap = 0;
    p = nullptr;
    return e;
  }
  cap = n;
  return hipSuccess;
}

template <typename CT, int D>
void launch_sort_passes(const KeyArgs& ka, PointSet& ps, int nbins, hipStream_t s) {
  const size_t lds = (size_t)nbins * sizeof(int32_t);
  int key_bits = 0;
  while ((1 << key_bits) < nbins) ++key_bits;
  hipLaunchKernelGGL((k_hist<CT, D>), dim3(ps.NB), dim3(kHistThreads), lds, s, ka, ps.M, ps.CH, ps.NB, nbins,
                     key_bits, ps.hist);
  const int64_t n = (int64_t)nbins * ps.NB;
  const int nparts = (int)((n + kScanTile - 1) / kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nparts), dim3(kScanThreads), 0, s, ps.hist, n, ps.scan_part);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, s, ps.scan_part, nparts);
  hipLaunchKernelGGL(k_scan_final, dim3(nparts), dim3(kScanThreads), 0, s, ps.hist, n, ps.scan_part);
  hipLaunchKernelGGL((k_scatter<CT, D>), dim3(ps.NB), dim3(64), lds, s, ka, ps.M, ps.CH, ps.NB, nbins, key_bits,
                     ps.hist, ps.perm);
}

}  // namespace

hipError_t pointset_sort(PointSet& ps, const Geometry& g, const SortTuning& tune, hipStream_t s, std::string& err) {
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
  int64_t CH = tune.chunk;
  int64_t NB = (M + CH - 1) / CH;
  if (NB * nbins > tune.hist_cap) {
    NB = std::max<int64_t>(1, tune.hist_cap / nbins);
    CH = (M + NB - 1) / NB;
    NB = (M + CH - 1) / CH;
  }
  if (NB < 1) NB = 1;
  ps.NB = (int)NB;
  ps.CH = (int)std::max<int64_t>(CH, 1);
  const int64_t nhist = (int64_t)nbins * NB;
  const int64_t nparts = (nhist + kScanTile - 1) / kScanTile;
  if (nparts > 1024 * 1024) {
    err = "histogram matrix too large";
    return hipErrorInvalidValue;
  }
  const int64_t max_sub = nbins + (M + ps.maxsub - 1) / ps.maxsub;
  if ((e = grow(ps.perm, ps.cap_M, M)) != hipSuccess) return e;
  if ((e = grow(ps.hist, ps.cap_hist, nhist)) != hipSuccess) return e;
  if ((e = grow(ps.scan_part, ps.cap_part, nparts)) != hipSuccess) return e;
  if ((e = grow(ps.bin_start, ps.cap_bins, (int64_t)nbins + 1)) != hipSuccess) return e;
  if ((e = grow(ps.subprob, ps.cap_sub, max_sub)) != hipSuccess) return e;
  if (!ps.nsub && (e = hipMalloc((void**)&ps.nsub, sizeof(int32_t))) != hipSuccess) return e;
  ps.max_subprob = (int)max_sub;

  if (M == 0) {
    hipLaunchKernelGGL(k_empty_sort, dim3(1), dim3(256), 0, s, nbins, ps.bin_start, ps.nsub);
    return hipGetLastError();
  }
  KeyArgs ka;
  for (int d = 0; d < 3; ++d) {
    ka.coord[d] = ps.coord[d];
    ka.nf[d] = g.nf[d];
    ka.bs[d] = g.bs[d];
    ka.nb[d] = g.nb[d];
  }
  if (ps.coord_f64) {
    if (g.dim == 1) launch_sort_passes<double, 1>(ka, ps, nbins, s);
    if (g.dim == 2) launch_sort_passes<double, 2>(ka, ps, nbins, s);
    if (g.dim == 3) launch_sort_passes<double, 3>(ka, ps, nbins, s);
  } else {
    if (g.dim == 1) launch_sort_passes<float, 1>(ka, ps, nbins, s);
    if (g.dim == 2) launch_sort_passes<float, 2>(ka, ps, nbins, s);
    if (g.dim == 3) launch_sort_passes<float, 3>(ka, ps, nbins, s);
  }
  hipLaunchKernelGGL(k_subprob, dim3(1), dim3(1024), 0, s, ps.hist, ps.NB, nbins, (int)M, ps.maxsub, ps.max_subprob,
                     ps.bin_start, ps.subprob, ps.nsub);
  return hipGetLastError();
}

void pointset_free(PointSet& ps) {
  void* bufs[] = {ps.perm, ps.hist, ps.scan_part, ps.bin_start, ps.subprob, ps.nsub};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  ps.perm = nullptr;
  ps.hist = nullptr;
  ps.scan_part = nullptr;
  ps.bin_start = nullptr;
  ps.subprob = nullptr;
  ps.nsub = nullptr;
  ps.cap_M = ps.cap_hist = ps.cap_part = ps.cap_bins = ps.cap_sub = 0;
}

}  // namespace nufftx
