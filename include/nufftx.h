/*
 * nufftx: MI355X (gfx950) implementation of nufftax's spread -> FFT -> deconvolve
 * pipeline behind a plain C ABI.
 *
 * Every pointer argument named as data (coordinates, strengths, modes, grids,
 * gradients) is a DEVICE pointer.  Complex data is interleaved (re, im) in the
 * plan precision: complex64 for NUFFTX_F32, complex128 for NUFFTX_F64;
 * coordinates are float32 / float64 respectively.  All work is enqueued on the
 * plan's stream; no call synchronises the host except makeplan/setpts when they
 * have to (re)allocate device buffers.  Every function returns a status code
 * (NUFFTX_OK on success); nufftx_last_error() gives the message of the last
 * failure on the calling thread.  Nothing aborts or exits.
 *
 * Layouts (nufftax conventions):
 *   strengths / point values  c[n_trans][M]
 *   modes  f[n_trans][n3][n2][n1]   (x fastest), CMCL order k in [-(n/2), (n-1)/2]
 *          when modeord == 0, FFT order when modeord == 1
 *   type-3 outputs            f[n_trans][N]
 *
 * Citations are to the reference Python package (GragasLab/nufftax).
 */
#ifndef NUFFTX_H
#define NUFFTX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NUFFTX_VERSION 100

enum nufftx_status {
  NUFFTX_OK = 0,
  NUFFTX_ERR_INVALID_ARG = 1,
  NUFFTX_ERR_ALLOC = 2,
  NUFFTX_ERR_HIP = 3,
  NUFFTX_ERR_FFT = 4,
  NUFFTX_ERR_STATE = 5
};

enum nufftx_precision { NUFFTX_F32 = 0, NUFFTX_F64 = 1 };

typedef struct nufftx_plan_s* nufftx_plan;

int nufftx_version(void);
const char* nufftx_last_error(void);

/* ---- host parameter helpers --------------------------------------------- */

/* Replaces nufftax.core.kernel.compute_kernel_params (nufftax/core/kernel.py:165-206). */
int nufftx_kernel_params(double eps, double upsampfac, int* nspread, double* beta, double* c);

/* Replaces nufftax.utils.grid.compute_grid_size (nufftax/utils/grid.py:36-58); <0 on error. */
int64_t nufftx_grid_size(int64_t n_modes, double upsampfac, int nspread);

/* ---- transforms ----------------------------------------------------------- */

/*
 * Plan one transform family.  Replaces the per-call setup inside
 * nufftax.transforms.nufft1.nufft{1,2,3}d1 (nufftax/transforms/nufft1.py:21-243),
 * nufft2.nufft{1,2,3}d2 (nufft2.py:21-227) and nufft3.nufft{1,2,3}d3
 * (nufft3.py:216-515): kernel parameters, fine-grid sizes, phihat / deconvolution
 * factors, FFT plans.
 *   type      1, 2 or 3
 *   dim       1, 2 or 3
 *   n_modes   type 1/2: mode counts (n1, n2, n3) -- x first;
 *             type 3:   fine-grid sizes (nf1, nf2, nf3), the reference's n_modes
 *   isign     sign of the exponent (+1 / -1)
 *   n_trans   number of transforms sharing the points (leading batch axis), 1..65535
 *   eps, upsampfac   as in the reference (static parameters)
 *   precision NUFFTX_F32 or NUFFTX_F64
 *   modeord   0 = CMCL (the public API), 1 = FFT order
 *   stream    hipStream_t (NULL = default stream)
 */
int nufftx_makeplan(int type, int dim, const int64_t* n_modes, int isign, int64_t n_trans, double eps,
                    double upsampfac, int precision, int modeord, void* stream, nufftx_plan* plan);

/* Fine-grid size of dimension d (0-based) of a plan; for type 3 the inner type-2 grid if inner != 0. */
int64_t nufftx_plan_grid_size(nufftx_plan plan, int d, int inner);

int nufftx_set_stream(nufftx_plan plan, void* stream);

/*
 * Set the nonuniform points: folds coordinates in FP64, bin-sorts them and builds
 * the subproblem list; the sort stays resident for every execute until the next
 * setpts.  Coordinates are read again by execute, so they must stay valid and
 * unchanged until then.  type 1/2: M points (x, y, z as dim requires);
 * type 3: M sources (x, y, z) and N targets (s, t, u); the type-3 centring,
 * rescaling and phase factors are computed here in FP64 (nufft3.py:257-299).
 */
int nufftx_setpts(nufftx_plan plan, int64_t M, const void* x, const void* y, const void* z, int64_t N,
                  const void* s, const void* t, const void* u);

/*
 * Run the transform.  type 1: c -> f; type 2: f -> c; type 3: c -> f.
 * Replaces the bodies of nufft{1,2,3}d1 (nufft1.py:51-89/126-162/201-243),
 * nufft{1,2,3}d2 (nufft2.py:48-85/114-152/183-227) and nufft{1,2,3}d3 (nufft3.py).
 */
int nufftx_execute(nufftx_plan plan, void* c, void* f);

/* Same as execute with the exponent sign overridden (type 1/2 only). */
int nufftx_execute_sign(nufftx_plan plan, void* c, void* f, int isign);

/*
 * Execute count plans (all type 1 or all type 2, each with its own points, grid size and
 * isign): plan i runs c[i] -> f[i] (type 1) or f[i] -> c[i] (type 2), as nufftx_execute.
 * No counterpart in the reference (its nearest relative is the n_trans batching of
 * nufftax/core/spread.py:100-173); for serving loops that batch requests of uneven M and
 * grid size.  Up to 8 FP32 2D/3D plans with the same kernel width, tile shape, n_trans and
 * stream share one spreading (type 1, the LDS-atomic kernel) or interpolation (type 2) launch;
 * other batches run plan by plan.  Shared launches equal the plans run one by one to FP32
 * rounding (relative l2 within 1e-6), not bit for bit.
 * The plans must be distinct.
 */
int nufftx_execute_many(int count, const nufftx_plan* plans, void* const* c, void* const* f);

/*
 * Reverse-mode derivative of a type-1 or type-2 plan (PyTorch cotangent
 * convention: cotangent = dL/dRe + i dL/dIm).  Replaces the JVP + transpose
 * rules of nufftax/transforms/primitives.py:88-110 (1D type 1), :127-151 (1D type 2),
 * :170-194 / :212-241 (2D), :260-288 / :306-340 (3D).
 *   type 1:  src = c[n_trans][M], g = cotangent of f  ->  grad_src = dL/dc (= type 2 with -isign of g),
 *            grad_x/y/z[M] = sum_t Re(conj(c_t) * sum_grid G_t dW/dx), G = FFT_{-isign}(pad(g / phihat))
 *   type 2:  src = f, g = cotangent of c[n_trans][M]  ->  grad_src = dL/df (= type 1 with -isign of g),
 *            grad_x/y/z[M] = sum_t Re(conj(g_t) * sum_grid F_t dW/dx), F = FFT_{isign}(pad(f / phihat))
 * Coordinate gradients use ES-kernel-derivative weights recomputed per point.
 * Any output pointer may be NULL to skip it.  grad_x/y/z are overwritten.
 */
int nufftx_vjp(nufftx_plan plan, const void* src, const void* g, void* grad_src, void* grad_x, void* grad_y,
               void* grad_z);

/* Frees a plan made by nufftx_makeplan / nufftx_makeplan_grid (not one from the plan cache). */
int nufftx_destroy(nufftx_plan plan);

/* ---- plan cache (for call sites that plan per call, such as an FFI handler) ---- */

/*
 * The reference plans nothing across calls: each primitive impl recomputes kernel parameters,
 * grid sizes and phihat (nufftax/transforms/primitives.py:64-70 lowers every call through
 * mlir.lower_fun(impl)).  A handler that replaces that lowering calls nufftx_plan_acquire with
 * the primitive's parameters and its stream, then nufftx_setpts + nufftx_execute(_sign), then
 * nufftx_plan_release.  Acquire returns an idle cached plan of the same parameters, stream and
 * device, or makes one (nufftx_makeplan's arguments and errors); the caller holds it alone until
 * release, so concurrent callers never share a plan.  Type-1/2 plans are shared between the
 * two signs: acquire sets the plan's sign to isign.  Released plans keep their buffers: a
 * setpts + execute of the same sizes on a reused plan allocates no device memory and does not
 * synchronise the host.  Release does not wait for the GPU (work in flight on a released plan
 * is ordered before the next holder's by the stream).  At most 32 idle plans are kept; beyond
 * that release destroys the least recently released one after synchronising its device.
 * Thread-safe.  A cached plan must not be passed to nufftx_destroy or nufftx_set_stream.  A
 * stream that cached plans were acquired for may be destroyed once they are released (an idle
 * plan is only handed out again for a stream with the same handle).
 */
int nufftx_plan_acquire(int type, int dim, const int64_t* n_modes, int isign, int64_t n_trans, double eps,
                        double upsampfac, int precision, int modeord, void* stream, nufftx_plan* out);

/* nufftx_plan_acquire for grid-operator plans (nufftx_makeplan_grid's arguments). */
int nufftx_plan_acquire_grid(int dim, const int64_t* nf, int nspread, double beta, double kc, int64_t n_trans,
                             int precision, void* stream, nufftx_plan* out);

/* Return an acquired plan to the cache. */
int nufftx_plan_release(nufftx_plan plan);

/* Destroy every idle cached plan (held plans stay), after synchronising their devices;
   returns how many were destroyed. */
int nufftx_plan_cache_clear(void);

/* Idle and held plans in the cache, plans made by acquire and acquires served from the cache. */
int nufftx_plan_cache_stats(int64_t* idle, int64_t* held, int64_t* made, int64_t* reused);

/* Device allocations (hipMalloc calls) the library has made so far in this process, rocFFT's
   own plan-time buffers aside: steady-state setpts/execute leave it unchanged. */
int64_t nufftx_device_allocations(void);

/* ---- stage-level operators (nufftax.core) ------------------------------- */

/*
 * Plan spreading/interpolation on an explicit fine grid nf (x first) with an
 * explicit ES kernel (nspread, beta, c): replaces the kernel_params argument of
 * nufftax.core.spread_{1,2,3}d / interp_{1,2,3}d (nufftax/core/spread.py:766-1305).
 * Points are set with nufftx_setpts (N ignored).
 */
int nufftx_makeplan_grid(int dim, const int64_t* nf, int nspread, double beta, double kc, int64_t n_trans,
                         int precision, void* stream, nufftx_plan* plan);

/* fw[n_trans][nf3][nf2][nf1] = spread(c)   (spread.py:285-339, 435-494, 596-658) */
int nufftx_spread(nufftx_plan plan, const void* c, void* fw);

/* fw += spread(c): accumulates into an existing grid (no clearing). */
int nufftx_spread_add(nufftx_plan plan, const void* c, void* fw);

/* c[n_trans][M] = interp(fw)                (spread.py:342-381, 497-540, 661-709) */
int nufftx_interp(nufftx_plan plan, const void* fw, void* c);

/*
 * Gather with kernel-derivative weights (spread.py:820-944, 997-1114, 1168-1305):
 *   gval[t][j] = sum G_t W_j                         (may be NULL)
 *   grad_ax[j] = sum_t Re(conj(P_t[j]) * sum G_t dW_j/d ax)   (each may be NULL; overwritten)
 * With P = c and G = grid cotangent this is the coordinate gradient of spread;
 * with P = point cotangent and G = fw it is the coordinate gradient of interp.
 */
int nufftx_gather_grad(nufftx_plan plan, const void* G, const void* P, void* gval, void* grad_x, void* grad_y,
                       void* grad_z);

/*
 * phihat[k], k = 0 .. nf/2, of the ES kernel (nspread, beta, c) by the FP64 trapezoid
 * rule the plans use (kernel_fourier_series, nufftax/core/kernel.py:209-260).
 * Host memory, out has nf/2 + 1 entries.
 */
int nufftx_kernel_fourier_series(int64_t nf, int nspread, double beta, double c, double* out);

/*
 * Deconvolution with mode truncation/padding on device buffers, enqueued on stream:
 *   direction 0: out[t][modes] = in[t][fine grid] * prod_d (-1)^k_d / phihat_d[|k_d|]
 *                (deconvolve_shuffle_{1,2,3}d, nufftax/core/deconvolve.py:81-111, 174-224, 312-366)
 *   direction 1: out[t][fine grid] = zero grid with the deconvolved in[t][modes] in FFT order
 *                (deconvolve_pad_{1,2,3}d, deconvolve.py:114-166, 227-304, 369-468)
 * phihat_d: device arrays of at least n_modes[d]/2 + 1 reals of the given precision.
 * Mode arrays are (n_trans, n3, n2, n1), grids (n_trans, nf3, nf2, nf1), x fastest.
 */
int nufftx_deconvolve(int direction, int dim, const int64_t* n_modes, const int64_t* nf, int64_t n_trans,
                      const void* phihat1, const void* phihat2, const void* phihat3, int modeord, int precision,
                      void* stream, const void* in, void* out);

/* Bin geometry of a plan (inner != 0: the inner grid of a type-3 plan). */
int nufftx_get_sort_info(nufftx_plan plan, int inner, int* nbins, int* bin_size /*[3]*/,
                         int* n_bins_per_dim /*[3]*/);

/* Copy the last setpts' sorted permutation (M or N int32) and bin starts
   (nbins + 1 int32) into device buffers, on the plan stream.  Either may be NULL. */
int nufftx_copy_sort(nufftx_plan plan, int inner, int32_t* perm, int32_t* bin_start);

/* ---- spreading without device allocation ---------------------------------------- */

/* Working memory of nufftx_spread_noalloc: a fixed device array of this many bytes per GPU,
   held from the library's first use on that GPU until the process ends. */
#define NUFFTX_NOALLOC_SCRATCH_BYTES (16 << 20)
/* Largest M nufftx_spread_noalloc accepts. */
#define NUFFTX_NOALLOC_MAX_POINTS (1 << 18)

/*
 * No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the result of
 * nufftx_makeplan_grid + nufftx_setpts + nufftx_spread for one transform (n_trans = 1),
 *   fw[nf3][nf2][nf1] = spread(c[M])   (spread.py:285-339, 435-494, 596-658),
 * with every working buffer (sort state, packed points, strength bounds) in a fixed device array
 * of NUFFTX_NOALLOC_SCRATCH_BYTES instead of memory allocated per call; no hipMalloc, hipFree or
 * allocating library call.  Inputs and output are the caller's device buffers; fw is overwritten.
 * Enqueued on stream.  Limits: M <= NUFFTX_NOALLOC_MAX_POINTS, and the sort state of the grid
 * must fit the array (NUFFTX_ERR_INVALID_ARG otherwise).  Calls that overlap on one GPU share
 * the array and race: serialise them (one stream, or events).
 */
int nufftx_spread_noalloc(int dim, const int64_t* nf, int nspread, double beta, double c, int precision, int64_t M,
                          const void* x, const void* y, const void* z, const void* cstr, void* fw, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NUFFTX_H */
