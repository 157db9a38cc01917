// Type-1 spreading launchers (kernels in si_kernels.hpp).
#include <cstdlib>

#include "si_mfma.hpp"

namespace nufftx {

namespace {

template <typename T, typename CT, int D>
hipError_t spread_impl(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  const SIArgs<T> a = make_si_args<T>(ga);
  size_t lds = 0;
  const bool tiled = use_tiles(*ga.g, ga.prec, lds, true);
  const PointSet& ps = *ga.ps;
  return dispatch_ns(ga.g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    if (tiled) {
      auto kern = k_spread_tiled<T, CT, D, NS>;
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

template <typename CT>
hipError_t spread_mfma(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  const SIArgs<float> a = make_si_args<float>(ga);
  const PointSet& ps = *ga.ps;
  auto launch = [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    const size_t lds = MfmaGeom<NS>::lds_bytes();
    hipLaunchKernelGGL((k_spread_mfma2d<CT, NS>), dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kSpreadThreads), lds,
                       s, a, static_cast<const cplx<float>*>(c), static_cast<cplx<float>*>(fw));
    return hipGetLastError();
  };
  switch (ga.g->ns) {
    case 2: return launch(std::integral_constant<int, 2>{});
    case 4: return launch(std::integral_constant<int, 4>{});
    case 6: return launch(std::integral_constant<int, 6>{});
    case 8: return launch(std::integral_constant<int, 8>{});
  }
  return hipErrorInvalidValue;
}

template <typename T, typename CT>
hipError_t spread_dim(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    if (mfma2d_applicable(*ga.g, ga.prec) && use_mfma_spread())
      return spread_mfma<CT>(ga, c, fw, s);
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
