// Gather with ES-kernel-derivative weights: coordinate gradients of spread and
// interp (nufftax/core/spread.py:820-1305) and the fused value+coordinate
// backward of the type-1/type-2 transforms.  Kernels in si_kernels.hpp.
#include <cstdlib>

#include "si_kernels.hpp"

namespace nufftx {

namespace {

template <typename T, typename CT, int D>
hipError_t grad_impl(const GridOpArgs& ga, const void* G, const void* P, void* gval, void* const gcoord[3],
                     hipStream_t s) {
  const SIArgs<T> a = make_si_args<T>(ga);
  size_t lds = 0;
  const bool tiled = use_tiles(*ga.g, ga.prec, lds);
  const PointSet& ps = *ga.ps;
  GatherOut<T, D, true> o;
  o.val = static_cast<cplx<T>*>(gval);
  o.P = static_cast<const cplx<T>*>(P);
  for (int d = 0; d < 3; ++d) o.gcoord[d] = d < D ? static_cast<T*>(gcoord[d]) : nullptr;
  return dispatch_ns(ga.g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    if (tiled) {
      if constexpr (std::is_same<T, float>::value && D >= 2 && NS <= 8) {
        // 16-byte tile reads, as in the value interpolation (NUFFTX_INTERP_WIDE=0: 8-byte)
        const char* ev = std::getenv("NUFFTX_INTERP_WIDE");
        const size_t lw = lds + 2 * sizeof(cplx<float>);  // the tile plus its two zero cells
        if ((a.tile[0] % 2) == 0 && lw <= (size_t)lds_device_max_bytes() && !(ev && std::atoi(ev) == 0)) {
          auto kw = k_gather_wide_deriv<CT, D, NS>;
          if (lw > 65536) (void)hipFuncSetAttribute((const void*)kw, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw);
          hipLaunchKernelGGL(kw, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(256), lw, s, a,
                             static_cast<const cplx<float>*>(G), o);
          return hipGetLastError();
        }
      }
      auto kern = k_gather_tiled<T, CT, D, NS, true>;
      if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(256), lds, s, a,
                         static_cast<const cplx<T>*>(G), o);
    } else {
      const unsigned nblk = (unsigned)((ps.M + 255) / 256);
      hipLaunchKernelGGL((k_gather_global<T, CT, D, NS, true>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s,
                         a, static_cast<const cplx<T>*>(G), o);
    }
    return hipGetLastError();
  });
}

template <typename T, typename CT>
hipError_t grad_dim(const GridOpArgs& ga, const void* G, const void* P, void* gval, void* const gcoord[3],
                    hipStream_t s) {
  switch (ga.g->dim) {
    case 1: return grad_impl<T, CT, 1>(ga, G, P, gval, gcoord, s);
    case 2: return grad_impl<T, CT, 2>(ga, G, P, gval, gcoord, s);
    case 3: return grad_impl<T, CT, 3>(ga, G, P, gval, gcoord, s);
  }
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_interp_grad(const GridOpArgs& ga, const void* G, const void* P, void* gval, void* const gcoord[3],
                              hipStream_t s) {
  if (ga.ps->M == 0) return hipSuccess;
  if (ga.prec == kF32) {
    return ga.ps->coord_f64 ? grad_dim<float, double>(ga, G, P, gval, gcoord, s)
                            : grad_dim<float, float>(ga, G, P, gval, gcoord, s);
  }
  return grad_dim<double, double>(ga, G, P, gval, gcoord, s);
}

}  // namespace nufftx
