// Setpts-time sub-bin order of the point records of the 2D FP32 matrix-core kernels.
//
// The bin-sort's scatter writes, at every sorted position, the record
//   {t_x, t_y, meta, p}:  t = i0 - xs (FP32, from the FP64 fold; i0 = ceil(xs - ns/2),
//   nufftax/core/spread.py:216-230), meta = ox | oy << 12 | s << 24 with (ox, oy) the
//   support start in the bin's tile and s the sub-bin, p the point.
// k_subsort then re-orders every subproblem's slice of records by sub-bin (stable:
// bin-sort order within a sub-bin), reading and writing only that slice, and
// records sub_off[sub][s], the first record of sub-bin s (sub_off[sub][16] = count).
// perm keeps the plain bin order.
#include "si_mfma.hpp"

namespace nufftx {

namespace {

constexpr int kPrepThreads = 256;
constexpr int kPrepPPT = 8;  // points per thread: subproblems of up to 2048 points

struct PrepArgs {
  const int4* subprob;
  const int32_t* nsub;
  float4* recs;
  int32_t* sub_off;
};

// Block-wide exclusive scan of 8 words of packed 16-bit counters.
__device__ __forceinline__ void block_scan_u16x8(const uint32_t (&v)[8], uint32_t (&excl)[8], uint32_t (&tot)[8]) {
  __shared__ uint32_t wsum[kPrepThreads / 64][8];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  uint32_t incl[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    incl[k] = v[k];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(incl[k], off, 64);
      if (lane >= off) incl[k] += t;
    }
  }
  if (lane == 63)
#pragma unroll
    for (int k = 0; k < 8; ++k) wsum[w][k] = incl[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t base = 0, all = 0;
#pragma unroll
    for (int ww = 0; ww < kPrepThreads / 64; ++ww) {
      const uint32_t s = wsum[ww][k];
      if (ww < w) base += s;
      all += s;
    }
    excl[k] = base + incl[k] - v[k];
    tot[k] = all;
  }
  __syncthreads();
}

__device__ __forceinline__ int field16(const uint32_t (&v)[8], int s) { return (int)((v[s >> 1] >> ((s & 1) * 16)) & 0xffff); }

// Records are loaded coalesced (record j = u * 256 + tid); the stable ranking walks
// the keys thread-contiguously (thread t owns records 8t .. 8t+7) through LDS.
__global__ void __launch_bounds__(kPrepThreads) k_subsort(PrepArgs a) {
  constexpr int CAP = kPrepThreads * kPrepPPT;
  __shared__ unsigned char skey[CAP];  // sub-bin of record j
  __shared__ int16_t spos[CAP];        // its position in sub-bin order
  const int sub = blockIdx.x;
  if (sub >= *a.nsub) return;
  const int4 sd = a.subprob[sub];
  const int tid = threadIdx.x;
  float4 r[kPrepPPT];
#pragma unroll
  for (int u = 0; u < kPrepPPT; ++u) {
    const int j = u * kPrepThreads + tid;
    if (j < sd.z) {
      r[u] = a.recs[sd.y + j];
      skey[j] = (unsigned char)((__float_as_int(r[u].z) >> 24) & 15);
    }
  }
  __syncthreads();
  int key[kPrepPPT];
  uint32_t cw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < kPrepPPT; ++u) {
    const int j = tid * kPrepPPT + u;
    key[u] = j < sd.z ? skey[j] : 16;
    if (key[u] < 16) cw[key[u] >> 1] += 1u << ((key[u] & 1) * 16);
  }
  uint32_t excl[8], tot[8];
  block_scan_u16x8(cw, excl, tot);
  int start[17];
  start[0] = 0;
#pragma unroll
  for (int s = 0; s < 16; ++s) start[s + 1] = start[s] + field16(tot, s);
  uint32_t seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int u = 0; u < kPrepPPT; ++u) {
    const int s = key[u];
    if (s >= 16) continue;
    spos[tid * kPrepPPT + u] = (int16_t)(start[s] + field16(excl, s) + field16(seen, s));
    seen[s >> 1] += 1u << ((s & 1) * 16);
  }
  __syncthreads();  // every record of the slice is in registers before the first write
#pragma unroll
  for (int u = 0; u < kPrepPPT; ++u) {
    const int j = u * kPrepThreads + tid;
    if (j < sd.z) a.recs[sd.y + spos[j]] = r[u];
  }
  if (tid <= 16) a.sub_off[(int64_t)sub * 17 + tid] = start[tid];
}

}  // namespace

bool subsort_applicable(const SortTuning& tune, const Geometry& g, int prec) {
  return use_mfma() && mfma2d_applicable(g, prec) && tune.maxsub <= kPrepThreads * kPrepPPT;
}

hipError_t pointset_subsort(PointSet& ps, const Geometry& g, hipStream_t s) {
  (void)g;
  ps.recs_subbin = false;
  if (ps.M == 0 || !ps.has_recs) return hipSuccess;
  hipError_t e;
  if ((e = grow_buffer(ps.sub_off, ps.cap_sub_off, (int64_t)ps.max_subprob * 17, ps.retired)) != hipSuccess) return e;
  PrepArgs a;
  a.subprob = ps.subprob;
  a.nsub = ps.nsub;
  a.recs = ps.recs;
  a.sub_off = ps.sub_off;
  hipLaunchKernelGGL(k_subsort, dim3((unsigned)std::max(1, ps.max_subprob)), dim3(kPrepThreads), 0, s, a);
  e = hipGetLastError();
  if (e == hipSuccess) ps.recs_subbin = true;
  return e;
}

}  // namespace nufftx
