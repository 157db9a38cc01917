// Type-1 spreading launchers (kernels in si_kernels.hpp).
#include <algorithm>
#include <cstdlib>

#include "si_mfma.hpp"

namespace nufftx {

namespace {

// cmax[t] = max |re|, |im| of the strengths of transform t (FP32 fixed-point scale).
hipError_t strength_absmax(const GridOpArgs& ga, const void* c, hipStream_t s) {
  PointSet& ps = const_cast<PointSet&>(*ga.ps);
  hipError_t e;
  const int nb = (int)std::min<int64_t>((ps.M + 1023) / 1024, kAbsmaxBlocks);
  // cmax holds n_trans floats followed by the per-block partial maxima
  if ((e = grow_buffer(ps.cmax, ps.cap_cmax, ga.n_trans * (1 + (int64_t)nb), ps.retired)) != hipSuccess) return e;
  unsigned* part = reinterpret_cast<unsigned*>(ps.cmax + ga.n_trans);
  hipLaunchKernelGGL(k_absmax_part, dim3(nb, (unsigned)ga.n_trans), dim3(256), 0, s, static_cast<const cplx<float>*>(c),
                     ps.M, part);
  hipLaunchKernelGGL(k_absmax_final, dim3((unsigned)ga.n_trans), dim3(256), 0, s, part, nb, ps.cmax);
  return hipGetLastError();
}

template <typename T, typename CT, int D>
hipError_t spread_impl(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  size_t lds = 0;
  const bool tiled = use_tiles(*ga.g, ga.prec, lds, true);
  const PointSet& ps = *ga.ps;
  if constexpr (std::is_same<T, float>::value) {
    if (tiled) {
      hipError_t e = strength_absmax(ga, c, s);
      if (e != hipSuccess) return e;
    }
  }
  SIArgs<T> a = make_si_args<T>(ga);
  a.cmax = ps.cmax;
  return dispatch_ns(ga.g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    if (tiled) {
      if constexpr (D >= 2 && std::is_same<T, float>::value) {
        // 2D/3D FP32: the bank-spread kernel when its point stage fits (maxsub <= 65535 points,
        // supports within 127 cells of the tile origin).  Tiles that leave room for one block
        // per CU get 1024 threads, the others 256 (several blocks per CU).
        const size_t lb = banked_lds<D>(a.tile_cells, ps.maxsub);
        bool fits = ps.maxsub <= 65535 && lb <= (size_t)lds_device_max_bytes();
        for (int d = 0; d < D; ++d) fits = fits && a.bs[d] <= 127;
        if (fits) {
          auto kb = k_spread_banked<CT, D, NS, false, kBankClasses>;
          if constexpr (D == 2)
            if (a.crec) kb = k_spread_banked<CT, D, NS, true, kBankClasses>;
          if (lb > 65536) (void)hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
          const unsigned nt = 2 * lb > (size_t)lds_device_max_bytes() ? 1024 : 256;
          hipLaunchKernelGGL(kb, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(nt), lb, s, a, ps.maxsub,
                             static_cast<const cplx<float>*>(c), static_cast<cplx<float>*>(fw));
          return hipGetLastError();
        }
      }
      auto kern = k_spread_tiled<T, CT, D, NS>;
      if constexpr (D == 2 && std::is_same<T, float>::value)
        if (a.crec) kern = k_spread_tiled<T, CT, D, NS, true>;
      if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(256), lds, s, a,
                         static_cast<const cplx<T>*>(c), static_cast<cplx<T>*>(fw));
    } else {
      const unsigned nblk = (unsigned)((ps.M + 255) / 256);
      hipLaunchKernelGGL((k_spread_global<T, CT, D, NS>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s, a,
                         static_cast<const cplx<T>*>(c), static_cast<cplx<T>*>(fw));
    }
    return hipGetLastError();
  });
}

// The records-based matrix-core spread, when the geometry has point records and the
// kernel's wave tiles fit in LDS; false means "use the LDS-tiled kernel".
bool spread_mfma(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s, hipError_t& err) {
  PointSet& ps = const_cast<PointSet&>(*ga.ps);
  if (!ps.recs_subbin) return false;
  const SIArgs<float> a = make_si_args<float>(ga);
  auto launch = [&](auto nsc) -> bool {
    constexpr int NS = decltype(nsc)::value;
    const size_t lds = MfmaGeom<NS>::spread_lds();
    if (lds > (size_t)lds_device_max_bytes()) return false;
    if ((err = strength_absmax(ga, c, s)) != hipSuccess) return true;
    auto kern = k_spread_mfma2d<NS>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kMfmaThreads), lds, s, a, ps.recs,
                       ps.sub_off, ps.cmax, static_cast<const cplx<float>*>(c), static_cast<cplx<float>*>(fw));
    err = hipGetLastError();
    return true;
  };
  switch (ga.g->ns) {
    case 2: return launch(std::integral_constant<int, 2>{});
    case 4: return launch(std::integral_constant<int, 4>{});
    case 6: return launch(std::integral_constant<int, 6>{});
    case 8: return launch(std::integral_constant<int, 8>{});
  }
  return false;
}

template <typename T, typename CT>
hipError_t spread_dim(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    hipError_t err = hipSuccess;
    if (mfma2d_applicable(*ga.g, ga.prec) && use_mfma() && spread_mfma(ga, c, fw, s, err)) return err;
  }
  switch (ga.g->dim) {
    case 1: return spread_impl<T, CT, 1>(ga, c, fw, s);
    case 2: return spread_impl<T, CT, 2>(ga, c, fw, s);
    case 3: return spread_impl<T, CT, 3>(ga, c, fw, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_spread(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  if (ga.ps->M == 0) return hipSuccess;
  if (ga.prec == kF32) {
    return ga.ps->coord_f64 ? spread_dim<float, double>(ga, c, fw, s) : spread_dim<float, float>(ga, c, fw, s);
  }
  return spread_dim<double, double>(ga, c, fw, s);
}

}  // namespace nufftx
