// Type-2 interpolation launchers (kernels in si_kernels.hpp and si_mfma.hpp).
#include <algorithm>
#include <cstdlib>

#include "si_mfma.hpp"
#include "si_win.hpp"

namespace nufftx {

namespace {

template <typename T, typename CT, int D>
hipError_t interp_impl(const GridOpArgs& ga, const void* fw, void* c, hipStream_t s) {
  const SIArgs<T> a = make_si_args<T>(ga);
  size_t lds = 0;
  const bool tiled = use_tiles(*ga.g, ga.prec, lds);
  const PointSet& ps = *ga.ps;
  GatherOut<T, D, false> o;
  o.val = static_cast<cplx<T>*>(c);
  o.P = nullptr;
  o.gcoord[0] = o.gcoord[1] = o.gcoord[2] = nullptr;
  return dispatch_ns(ga.g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    if (tiled) {
      if constexpr (D >= 2 && std::is_same<T, float>::value && NS <= 8) {
        // 16-byte tile reads (k_gather_wide) unless NUFFTX_INTERP_WIDE=0; rows of an even
        // number of cells keep the reads aligned
        const char* ev = std::getenv("NUFFTX_INTERP_WIDE");
        const size_t lw = lds + 2 * sizeof(cplx<float>);  // the tile plus its two zero cells
        if ((a.tile[0] % 2) == 0 && lw <= (size_t)lds_device_max_bytes() && !(ev && std::atoi(ev) == 0)) {
          auto kw = k_gather_wide<CT, D, NS, false>;
          if constexpr (D == 2)
            if (a.crec) kw = k_gather_wide<CT, D, NS, true>;
          if constexpr (NS == kEsPolyNs)
            if (a.poly.on) {
              kw = k_gather_wide<CT, D, NS, false, true>;
              if constexpr (D == 2)
                if (a.crec) kw = k_gather_wide<CT, D, NS, true, true>;
            }
          if (lw > 65536) (void)hipFuncSetAttribute((const void*)kw, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lw);
          hipLaunchKernelGGL(kw, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(256), lw, s, a,
                             static_cast<const cplx<float>*>(fw), static_cast<cplx<float>*>(c));
          return hipGetLastError();
        }
      }
      auto kern = k_gather_tiled<T, CT, D, NS, false>;
      if constexpr (D == 2 && std::is_same<T, float>::value)
        if (a.crec) kern = k_gather_tiled<T, CT, D, NS, false, true>;
      if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(256), lds, s, a,
                         static_cast<const cplx<T>*>(fw), o);
    } else {
      const char* gw = std::getenv("NUFFTX_GLOBAL_WAVE");
      if (D >= 2 && !(gw && gw[0] == '0')) {
        // one wave per point (four per block)
        const unsigned nblk = (unsigned)((ps.M + 3) / 4);
        hipLaunchKernelGGL((k_gather_global_wave<T, CT, D == 1 ? 2 : D, NS>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s, a,
                           static_cast<const cplx<T>*>(fw), static_cast<cplx<T>*>(c));
      } else {
        const unsigned nblk = (unsigned)((ps.M + 255) / 256);
        hipLaunchKernelGGL((k_gather_global<T, CT, D, NS, false>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s,
                           a, static_cast<const cplx<T>*>(fw), o);
      }
    }
    return hipGetLastError();
  });
}

// The records-based matrix-core interpolation (needs the setpts point records).
bool interp_mfma(const GridOpArgs& ga, const void* fw, void* c, hipStream_t s, hipError_t& err) {
  const PointSet& ps = *ga.ps;
  if (!ps.recs_subbin) return false;
  const SIArgs<float> a = make_si_args<float>(ga);
  auto launch = [&](auto nsc) -> bool {
    constexpr int NS = decltype(nsc)::value;
    const size_t lds = MfmaGeom<NS>::interp_lds();
    if (lds > (size_t)lds_device_max_bytes()) return false;
    auto kern = k_interp_mfma2d<NS>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kMfmaThreads), lds, s, a, ps.recs,
                       ps.sub_off, static_cast<const cplx<float>*>(fw), static_cast<cplx<float>*>(c));
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

// The matrix-core window interpolation (2D FP32, ns <= 8, window bins, compact records).
bool interp_win(const GridOpArgs& ga, const void* fw, void* c, hipStream_t s, hipError_t& err) {
  const PointSet& ps = *ga.ps;
  if (!win2d_geometry(*ga.g, ga.prec, ps.maxsub) || !ps.has_crec || !win_interp_enabled()) return false;
  const SIArgs<float> a = make_si_args<float>(ga);
  if (!a.crec) return false;
  auto launch = [&](auto nsc) -> bool {
    constexpr int NS = decltype(nsc)::value;
    const size_t lds = wi_interp_lds<NS>(ps.maxsub);
    // static LDS (ordering counters) below 2.2 KB
    if (lds + 2304 > (size_t)lds_device_max_bytes()) return false;
    auto kern = k_interp_win2dh<NS>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kWiThreads), lds, s, a, ps.maxsub,
                       static_cast<const cplx<float>*>(fw), static_cast<cplx<float>*>(c));
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
hipError_t interp_dim(const GridOpArgs& ga, const void* fw, void* c, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    hipError_t err = hipSuccess;
    if (mfma2d_applicable(*ga.g, ga.prec) && use_mfma() && interp_mfma(ga, fw, c, s, err)) return err;
    if (interp_win(ga, fw, c, s, err)) return err;
  }
  switch (ga.g->dim) {
    case 1: return interp_impl<T, CT, 1>(ga, fw, c, s);
    case 2: return interp_impl<T, CT, 2>(ga, fw, c, s);
    case 3: return interp_impl<T, CT, 3>(ga, fw, c, s);
  }
  return hipErrorInvalidValue;
}

template <typename CT, int D>
hipError_t interp_many_impl(const GridOpArgs* gas, int count, const void* const* fw, void* const* c, hipStream_t s) {
  size_t lmax = 0;
  GatherMany m;
  m.count = count;
  m.first[0] = 0;
  for (int i = 0; i < count; ++i) {
    size_t l = 0;
    if (!use_tiles(*gas[i].g, gas[i].prec, l)) return hipErrorNotSupported;
    lmax = std::max(lmax, l);
    m.a[i] = make_si_args<float>(gas[i]);
    m.fw[i] = static_cast<const cplx<float>*>(fw[i]);
    m.out[i] = static_cast<cplx<float>*>(c[i]);
    m.first[i + 1] = m.first[i] + std::max(0, gas[i].ps->max_subprob);
  }
  for (int i = count; i < kMaxMany; ++i) m.first[i + 1] = m.first[count];
  if (m.first[count] == 0) return hipSuccess;
  const bool rec = m.a[0].crec != nullptr;
  for (int i = 1; i < count; ++i)
    if ((m.a[i].crec != nullptr) != rec) return hipErrorNotSupported;
  return dispatch_ns(gas[0].g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    auto kern = k_gather_tiled_many<CT, D, NS, false>;
    if constexpr (D == 2)
      if (rec) kern = k_gather_tiled_many<CT, D, NS, true>;
    if (lmax > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lmax);
    hipLaunchKernelGGL(kern, dim3((unsigned)m.first[count], (unsigned)gas[0].n_trans), dim3(256), lmax, s, m);
    return hipGetLastError();
  });
}

}  // namespace

hipError_t launch_interp_many(const GridOpArgs* gas, int count, const void* const* fw, void* const* c, hipStream_t s) {
  if (count < 1 || count > kMaxMany) return hipErrorNotSupported;
  const GridOpArgs& g0 = gas[0];
  if (g0.prec != kF32 || g0.g->dim < 2 || use_mfma()) return hipErrorNotSupported;
  for (int i = 0; i < count; ++i) {
    const GridOpArgs& g = gas[i];
    if (g.prec != kF32 || g.g->dim != g0.g->dim || g.g->ns != g0.g->ns || g.n_trans != g0.n_trans ||
        g.ps->coord_f64 != g0.ps->coord_f64 || g.beta != g0.beta || g.c != g0.c || g.ps->M == 0 ||
        g.g->tile_cells != g0.g->tile_cells)  // one LDS size for the whole launch (see launch_spread_many)
      return hipErrorNotSupported;
  }
  const bool f64 = g0.ps->coord_f64;
  if (g0.g->dim == 2)
    return f64 ? interp_many_impl<double, 2>(gas, count, fw, c, s) : interp_many_impl<float, 2>(gas, count, fw, c, s);
  return f64 ? interp_many_impl<double, 3>(gas, count, fw, c, s) : interp_many_impl<float, 3>(gas, count, fw, c, s);
}

hipError_t launch_interp(const GridOpArgs& ga, const void* fw, void* c, hipStream_t s) {
  if (ga.ps->M == 0) return hipSuccess;
  if (ga.prec == kF32) {
    return ga.ps->coord_f64 ? interp_dim<float, double>(ga, fw, c, s) : interp_dim<float, float>(ga, fw, c, s);
  }
  return interp_dim<double, double>(ga, fw, c, s);
}

}  // namespace nufftx
