// Type-1 spreading launchers (kernels in si_kernels.hpp).
#include <algorithm>
#include <cstdlib>

#include "si_mfma.hpp"
#include "si_win.hpp"
#include "si_win3.hpp"

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
        // per CU get 1024 threads, the others 512 (several blocks per CU; measured on the 2D
        // radial workload: 211-214 us against 224-230 us with 256; NUFFTX_SPREAD_THREADS
        // overrides).
        // packed tiles (one atomic per support cell) for plans of eps >= 1e-4 with at most
        // 1024 points per subproblem, the bound the packed scale assumes
        const bool pk = ga.tol >= kPackedTol && ps.maxsub <= 1024 && !std::getenv("NUFFTX_NO_PACKED");
        const size_t lb = banked_lds<D>(a.tile_cells, ps.maxsub, pk);
        // the static class counters (2 x kBankClasses ints) share the block's LDS
        bool fits = ps.maxsub <= 65535 && lb + 2 * kBankClasses * sizeof(int) <= (size_t)lds_device_max_bytes();
        for (int d = 0; d < D; ++d) fits = fits && a.bs[d] <= 127;
        if (fits) {
          auto kb = pk ? k_spread_banked<CT, D, NS, false, kBankClasses, true>
                       : k_spread_banked<CT, D, NS, false, kBankClasses>;
          if constexpr (D == 2)
            if (a.crec)
              kb = pk ? k_spread_banked<CT, D, NS, true, kBankClasses, true> : k_spread_banked<CT, D, NS, true, kBankClasses>;
          if (lb > 65536) (void)hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lb);
          unsigned nt = 2 * lb > (size_t)lds_device_max_bytes() ? 1024 : 512;
          if (const char* e = std::getenv("NUFFTX_SPREAD_THREADS"))
            nt = (unsigned)std::min(1024, std::max(64, std::atoi(e) / 64 * 64));
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
      const char* gw = std::getenv("NUFFTX_GLOBAL_WAVE");
      if (D >= 2 && !(gw && gw[0] == '0')) {
        // one wave per point (four per block)
        const unsigned nblk = (unsigned)((ps.M + 3) / 4);
        hipLaunchKernelGGL((k_spread_global_wave<T, CT, D == 1 ? 2 : D, NS>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s, a,
                           static_cast<const cplx<T>*>(c), static_cast<cplx<T>*>(fw));
      } else {
        const unsigned nblk = (unsigned)((ps.M + 255) / 256);
        hipLaunchKernelGGL((k_spread_global<T, CT, D, NS>), dim3(nblk, (unsigned)ga.n_trans), dim3(256), 0, s, a,
                           static_cast<const cplx<T>*>(c), static_cast<cplx<T>*>(fw));
      }
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

// The matrix-core window spread (2D FP32, ns <= 8, subproblems of <= 4096 points): the
// points are packed in sorted order first (the sort left the inverse permutation), then
// spread from those contiguous records.
bool spread_win(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s, hipError_t& err) {
  PointSet& ps = const_cast<PointSet&>(*ga.ps);
  if (!win2d_geometry(*ga.g, ga.prec, ps.maxsub) || !ps.has_slot || !win_spread_enabled()) return false;
  SIArgs<float> a = make_si_args<float>(ga);
  // zero_out: the packing pass clears the grid (n_trans grids of complex64; an odd cell count
  // leaves 8 bytes over the pass's 16-byte stores, and takes the fill kernel)
  const int64_t grid_bytes = ga.n_trans * ga.g->grid_cells() * (int64_t)sizeof(cplx<float>);
  const bool fused_zero = ga.zero_out && grid_bytes % 16 == 0;
  void* zero = fused_zero ? fw : nullptr;
  const int64_t zero_bytes = fused_zero ? grid_bytes : 0;
  auto launch = [&](auto nsc) -> bool {
    constexpr int NS = decltype(nsc)::value;
    if (!win_f32_kernel() || ps.maxsub > 1024) {
      // the fp16-split kernel; its static LDS (counters) is below 2.2 KB
      const size_t lds = wh_spread_lds<NS>(ps.maxsub);
      if (lds + 2304 > (size_t)lds_device_max_bytes()) return false;
      if (ga.zero_out && !fused_zero && (err = launch_zero(fw, grid_bytes, s)) != hipSuccess) return true;
      if ((err = pointset_pack2d(ps, *ga.g, c, ga.n_trans, s, zero, zero_bytes)) != hipSuccess) return true;
      auto kern = k_spread_win2dh<NS, false>;
      if constexpr (NS == kEsPolyNs)
        if (a.poly.on) kern = k_spread_win2dh<NS, true>;
      if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
      hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kWhThreads), lds, s, a, ps.maxsub,
                         ps.packed, static_cast<cplx<float>*>(fw));
      err = hipGetLastError();
      return true;
    }
    const bool exact = win_exact_tiles();
    const size_t lds = exact ? win_spread_lds<NS, false>(ps.maxsub) : win_spread_lds<NS, true>(ps.maxsub);
    if (lds + 2048 > (size_t)lds_device_max_bytes()) return false;
    if (ga.zero_out && !fused_zero && (err = launch_zero(fw, grid_bytes, s)) != hipSuccess) return true;
    if ((err = pointset_pack2d(ps, *ga.g, c, ga.n_trans, s, zero, zero_bytes)) != hipSuccess) return true;
    auto kern = exact ? k_spread_win2d<NS, false> : k_spread_win2d<NS, true>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kWinThreads), lds, s, a, ps.maxsub,
                       ps.packed, static_cast<cplx<float>*>(fw));
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

// The 3D matrix-core window spread (3D FP32, ns = 8, bins of whole 8 x 8 x 4 sub-bins,
// subproblems of <= 2048 points): points packed in sorted order, then spread.  Exact tiles
// (two 64-bit words per cell: the bank-spread kernel's accuracy) are the default; packed
// tiles (one word per cell, about 3x the rounding error, 2% faster on 8M clustered points)
// are taken with NUFFTX_WIN3_PACKED=1 or where the exact tile does not fit the LDS.
bool spread_win3d(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s, hipError_t& err) {
  PointSet& ps = const_cast<PointSet&>(*ga.ps);
  if (!win3d_geometry(*ga.g, ga.prec, ps.maxsub) || !ps.has_slot || !win_spread_enabled()) return false;
  SIArgs<float> a = make_si_args<float>(ga);
  // the static LDS (ordering counters) is below 2 KB
  const size_t lds_exact = w3_spread_lds<false>(ga.g->tile_cells, ps.maxsub);
  const size_t lds_packed = w3_spread_lds<true>(ga.g->tile_cells, ps.maxsub);
  const char* pe = std::getenv("NUFFTX_WIN3_PACKED");
  const bool exact = lds_exact + 2048 <= (size_t)lds_device_max_bytes() && !(pe && pe[0] == '1');
  const size_t lds = exact ? lds_exact : lds_packed;
  if (lds + 2048 > (size_t)lds_device_max_bytes()) return false;
  const int64_t grid_bytes = ga.n_trans * ga.g->grid_cells() * (int64_t)sizeof(cplx<float>);
  const bool fused_zero = ga.zero_out && grid_bytes % 16 == 0;
  if (ga.zero_out && !fused_zero && (err = launch_zero(fw, grid_bytes, s)) != hipSuccess) return true;
  if ((err = pointset_pack3d(ps, *ga.g, c, ga.n_trans, s, fused_zero ? fw : nullptr, fused_zero ? grid_bytes : 0)) !=
      hipSuccess)
    return true;
  // ns = 8 (win3d_geometry); the exp form of the kernel weights (the polynomial form spilled
  // scalar registers here and was 0.035 ms slower on 8M clustered points)
  auto kern = exact ? k_spread_win3dh<8, false> : k_spread_win3dh<8, true>;
  if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(kern, dim3(ps.max_subprob, (unsigned)ga.n_trans), dim3(kW3Threads), lds, s, a, ps.maxsub,
                     ps.packed, ps.cs, static_cast<cplx<float>*>(fw));
  err = hipGetLastError();
  return true;
}

template <typename T, typename CT>
hipError_t spread_dim(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  if constexpr (std::is_same<T, float>::value) {
    hipError_t err = hipSuccess;
    const bool mf = mfma2d_applicable(*ga.g, ga.prec) && use_mfma();
    if (!mf && spread_win(ga, c, fw, s, err)) return err;  // clears fw itself when asked
    if (spread_win3d(ga, c, fw, s, err)) return err;       // likewise
    GridOpArgs gz = ga;
    gz.zero_out = false;
    if (ga.zero_out && (err = launch_zero(fw, ga.n_trans * ga.g->grid_cells() * (int64_t)sizeof(cplx<T>), s)) != hipSuccess)
      return err;
    if (mf && spread_mfma(gz, c, fw, s, err)) return err;
    if (mf && spread_win(gz, c, fw, s, err)) return err;
    switch (ga.g->dim) {
      case 1: return spread_impl<T, CT, 1>(gz, c, fw, s);
      case 2: return spread_impl<T, CT, 2>(gz, c, fw, s);
      case 3: return spread_impl<T, CT, 3>(gz, c, fw, s);
    }
    return hipErrorInvalidValue;
  }
  if (ga.zero_out) {
    hipError_t err = launch_zero(fw, ga.n_trans * ga.g->grid_cells() * (int64_t)sizeof(cplx<T>), s);
    if (err != hipSuccess) return err;
  }
  switch (ga.g->dim) {
    case 1: return spread_impl<T, CT, 1>(ga, c, fw, s);
    case 2: return spread_impl<T, CT, 2>(ga, c, fw, s);
    case 3: return spread_impl<T, CT, 3>(ga, c, fw, s);
  }
  return hipErrorInvalidValue;
}

// Whether plan ga can join a many-plan bank-spread launch; lds gets its LDS bytes.
template <int D>
bool banked_ok(const GridOpArgs& ga, size_t& lds) {
  size_t tl = 0;
  if (!use_tiles(*ga.g, ga.prec, tl, true)) return false;
  const PointSet& ps = *ga.ps;
  lds = banked_lds<D>(ga.g->tile_cells, ps.maxsub);
  if (ps.maxsub > 65535 || lds + 2 * kBankClasses * sizeof(int) > (size_t)lds_device_max_bytes()) return false;
  for (int d = 0; d < D; ++d)
    if (ga.g->bs[d] > 127) return false;
  return true;
}

template <typename CT, int D>
hipError_t spread_many_impl(const GridOpArgs* gas, int count, const void* const* c, void* const* fw, hipStream_t s) {
  size_t lmax = 0;
  for (int i = 0; i < count; ++i) {
    size_t l = 0;
    if (!banked_ok<D>(gas[i], l)) return hipErrorNotSupported;
    lmax = std::max(lmax, l);
  }
  SpreadMany m;
  m.count = count;
  m.maxsub = gas[0].ps->maxsub;
  m.first[0] = 0;
  for (int i = 0; i < count; ++i) {
    hipError_t e = strength_absmax(gas[i], c[i], s);
    if (e != hipSuccess) return e;
    m.a[i] = make_si_args<float>(gas[i]);
    m.a[i].cmax = gas[i].ps->cmax;
    m.c[i] = static_cast<const cplx<float>*>(c[i]);
    m.fw[i] = static_cast<cplx<float>*>(fw[i]);
    m.first[i + 1] = m.first[i] + std::max(0, gas[i].ps->max_subprob);
  }
  for (int i = count; i < kMaxMany; ++i) m.first[i + 1] = m.first[count];
  if (m.first[count] == 0) return hipSuccess;
  const bool rec = m.a[0].crec != nullptr;
  for (int i = 1; i < count; ++i)
    if ((m.a[i].crec != nullptr) != rec) return hipErrorNotSupported;
  const unsigned nt = 2 * lmax > (size_t)lds_device_max_bytes() ? 1024 : 256;
  return dispatch_ns(gas[0].g->ns, [&](auto nsc) -> hipError_t {
    constexpr int NS = decltype(nsc)::value;
    auto kb = k_spread_banked_many<CT, D, NS, false, kBankClasses>;
    if constexpr (D == 2)
      if (rec) kb = k_spread_banked_many<CT, D, NS, true, kBankClasses>;
    if (lmax > 65536) (void)hipFuncSetAttribute((const void*)kb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lmax);
    hipLaunchKernelGGL(kb, dim3((unsigned)m.first[count], (unsigned)gas[0].n_trans), dim3(nt), lmax, s, m);
    return hipGetLastError();
  });
}

}  // namespace

hipError_t launch_spread_many(const GridOpArgs* gas, int count, const void* const* c, void* const* fw, hipStream_t s) {
  if (count < 1 || count > kMaxMany) return hipErrorNotSupported;
  const GridOpArgs& g0 = gas[0];
  if (g0.prec != kF32 || g0.g->dim < 2 || use_mfma()) return hipErrorNotSupported;
  for (int i = 1; i < count; ++i) {
    const GridOpArgs& g = gas[i];
    if (g.prec != kF32 || g.g->dim != g0.g->dim || g.g->ns != g0.g->ns || g.n_trans != g0.n_trans ||
        g.ps->coord_f64 != g0.ps->coord_f64 || g.ps->maxsub != g0.ps->maxsub || g.beta != g0.beta || g.c != g0.c)
      return hipErrorNotSupported;
  }
  // a shared launch sizes its LDS for the largest tile; plans with other tile shapes would
  // run their blocks at that (lower) occupancy, so they launch separately
  for (int i = 0; i < count; ++i)
    if (gas[i].ps->M == 0 || gas[i].g->tile_cells != g0.g->tile_cells) return hipErrorNotSupported;
  const bool f64 = g0.ps->coord_f64;
  if (g0.g->dim == 2)
    return f64 ? spread_many_impl<double, 2>(gas, count, c, fw, s) : spread_many_impl<float, 2>(gas, count, c, fw, s);
  return f64 ? spread_many_impl<double, 3>(gas, count, c, fw, s) : spread_many_impl<float, 3>(gas, count, c, fw, s);
}

hipError_t launch_spread(const GridOpArgs& ga, const void* c, void* fw, hipStream_t s) {
  if (ga.ps->M == 0)
    return ga.zero_out ? launch_zero(fw, ga.n_trans * ga.g->grid_cells() * (ga.prec == kF32 ? 8 : 16), s) : hipSuccess;
  if (ga.prec == kF32) {
    return ga.ps->coord_f64 ? spread_dim<float, double>(ga, c, fw, s) : spread_dim<float, float>(ga, c, fw, s);
  }
  return spread_dim<double, double>(ga, c, fw, s);
}

}  // namespace nufftx
