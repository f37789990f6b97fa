/*
 * afft.h -- C ABI of the MI355X (gfx950) FFT engine behind Data.Array.Accelerate.Math.FFT.
 *
 * Everything here is plain C: device pointers are `void *`, sizes are int64_t, the HIP stream is
 * an opaque `void *` (a hipStream_t, NULL = the default stream).  Every function returns an
 * `int` status (AFFT_SUCCESS = 0, see afft_strerror) and nothing throws across this boundary.
 *
 * Layout contract (src/Data/Array/Accelerate/Math/FFT/Type.hs:28-39): an element of
 * `Complex Float` / `Complex Double` is an interleaved (re, im) pair (Vec2 e), 8 or 16 bytes;
 * arrays are row-major in Accelerate's shape order Z :. e_0 :. ... :. e_{rank-1}, the innermost
 * extent e_{rank-1} varying fastest.  `extents` arguments list e_0 .. e_{rank-1}.
 *
 * Semantics of `mode` (src/Data/Array/Accelerate/Math/FFT/Mode.hs:15-26):
 *   AFFT_FORWARD  sum_n x[n] exp(-2 pi i nk/N)
 *   AFFT_REVERSE  sum_n x[n] exp(+2 pi i nk/N), unnormalised
 *   AFFT_INVERSE  as Reverse, then divided by N, fused into the last pass.  N is the innermost
 *                 extent for afft_fft (FFT.hs:69), the length for afft_fft1D (FFT.hs:98) and the
 *                 total size for afft_fft2D/afft_fft3D (FFT.hs:125,155).
 *
 * All transforms are out-of-place (in-place when d_in == d_out is also allowed) and leave the
 * input unmodified when out-of-place, as the reference plugins do (LLVM/PTX.hs:92 writes a fresh
 * array; LLVM/Native/Base.hs:47 asks FFTW to preserve the input).  Calls are asynchronous on
 * the given stream, like the cuFFT path they replace (LLVM/PTX.hs:93,121-124).
 */
#ifndef AFFT_H
#define AFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------------- */
#define AFFT_SUCCESS 0
#define AFFT_ERROR_INVALID_VALUE 1    /* null pointer, negative extent, bad mode/axes value  */
#define AFFT_ERROR_UNSUPPORTED_RANK 2 /* rank outside the entry point's range                */
#define AFFT_ERROR_UNSUPPORTED_DTYPE 3
#define AFFT_ERROR_UNSUPPORTED_SIZE 4 /* a length the engine cannot plan                      */
#define AFFT_ERROR_ALLOC 5            /* device or host allocation failed                     */
#define AFFT_ERROR_HIP 6              /* a HIP runtime call failed                            */
#define AFFT_ERROR_NO_DEVICE 7        /* no usable gfx950 device                              */
#define AFFT_ERROR_IO 8               /* plan-cache file could not be read or written        */
#define AFFT_ERROR_INTERNAL 9
#define AFFT_ERROR_SCRATCH_TOO_SMALL 10 /* afft_exec_noalloc: the plan needs more than the scratch */

/* ---- enums ------------------------------------------------------------------------------ */
/* element type: NumericRfloat32 / NumericRfloat64 (FFT/Type.hs:28-30) */
#define AFFT_C64 0  /* Complex Float  (cuFFT C2C, LLVM/PTX.hs:126-127) */
#define AFFT_C128 1 /* Complex Double (cuFFT Z2Z, LLVM/PTX.hs:128)     */

/* Mode (FFT/Mode.hs:15-19) */
#define AFFT_FORWARD 0
#define AFFT_REVERSE 1
#define AFFT_INVERSE 2

/* which axes a plan transforms */
#define AFFT_AXES_INNERMOST 0 /* fft: innermost axis, outer axes are a batch */
#define AFFT_AXES_ALL 1       /* fft1D / fft2D / fft3D: every axis           */
#define AFFT_AXES_LAST2 2     /* batched fft2D: the two innermost axes       */

typedef struct afft_plan afft_plan;
typedef struct afft_ragged_plan afft_ragged_plan;

/* ---- library ---------------------------------------------------------------------------- */

/* Human-readable text for a status code; never NULL. */
const char *afft_strerror(int status);

/* Library version as major*10000 + minor*100 + patch. */
int afft_version(void);

/* ---- the four front-end operations (the plugin entry points) ----------------------------- */
/*
 * Each call plans on first use of a (device, dtype, shape, axes) key and reuses the plan from a
 * mutex-guarded cache keyed on the full tuple afterwards.
 */

/* fft: DFT along the innermost axis of an array of rank 1..8; the outer axes are a batch.
 * Replaces LLVM/PTX.hs:52-61 (cuFFT planMany over the innermost extent, rank <= 3) and
 * LLVM/Native.hs:49-72 (FFTW dftGU over axis [rank-1], rank <= 5); the front-end is FFT.hs:63-84. */
int afft_fft(int mode, int dtype, int rank, const int64_t *extents, const void *d_in, void *d_out,
             void *stream);

/* fft1D: DFT of a vector of length n.
 * Replaces LLVM/PTX.hs:63-64 (plan1D, :141) and LLVM/Native.hs:74-82; front-end FFT.hs:92-111. */
int afft_fft1D(int mode, int dtype, int64_t n, const void *d_in, void *d_out, void *stream);

/* fft2D: separable DFT over both axes of an h x w matrix (w innermost).
 * Replaces LLVM/PTX.hs:66-67 (plan2D h w, :148) and LLVM/Native.hs:84-92; front-end FFT.hs:119-142. */
int afft_fft2D(int mode, int dtype, int64_t h, int64_t w, const void *d_in, void *d_out,
               void *stream);

/* fft3D: DFT over all axes of a d x h x w array (w innermost).
 * Replaces LLVM/PTX.hs:69-70 (plan3D d h w, :155) and LLVM/Native.hs:94-102; front-end FFT.hs:150-173. */
int afft_fft3D(int mode, int dtype, int64_t d, int64_t h, int64_t w, const void *d_in, void *d_out,
               void *stream);

/* fft2D over each of `batch` h x w matrices stored back to back (the two innermost axes of a
 * batch x h x w array); Inverse divides by h*w.  Extension: the reference API has no batched 2D
 * call (fft2D takes a DIM2 only, FFT.hs:119-122), so a batch needs one fft2D per matrix. */
int afft_fft2D_many(int mode, int dtype, int64_t batch, int64_t h, int64_t w, const void *d_in,
                    void *d_out, void *stream);

/* Centring and shifting (DFT/Centre.hs:36-164) of a rank 1-3 array of fewer than 2^31 elements.
 * `which`: AFFT_CENTRE  x * (-1)^(sum of the multi-index)                 (centre1D/2D/3D, :36-65)
 *          AFFT_SHIFT   y[i] = x[(i + ceil(n/2)) mod n] along every axis   (shift1D/2D/3D)
 *          AFFT_ISHIFT  y[i] = x[(i + floor(n/2)) mod n] along every axis  (ishift1D/2D/3D)
 * Extension: the reference computes these as separate Accelerate arrays. */
#define AFFT_REORDER_NONE 0
#define AFFT_CENTRE 1
#define AFFT_SHIFT 2
#define AFFT_ISHIFT 3
int afft_reorder(int which, int dtype, int rank, const int64_t *extents, const void *d_in, void *d_out,
                 void *stream);
/* post(fftND(mode, pre(x))) over all axes of a rank 1-3 array (fft1D/2D/3D of FFT.hs:92-173 with
 * the Centre.hs operations on either side), pre / post as `which` above.  A centring, and a shift
 * or ishift when every extent is even, is folded into the transform's first load or last store
 * (no extra pass: shift(F x) = F(x (-1)^m) and F(shift x) = (-1)^k F x for even n); a shift over
 * an odd extent adds one reordering pass. */
int afft_fft_centred(int mode, int dtype, int rank, const int64_t *extents, int pre, int post,
                     const void *d_in, void *d_out, void *stream);

/* Data.Array.Accelerate.Math.DFT (DFT.hs:42-111), O(N^2) on the device, for validation and for
 * extents the FFT would pad.  As the reference's dftG, the sum runs over EVERY element of the array
 * (DFT.hs:68-92): out[ix] = sum_i in[i] * roots[sh_i :. (k * n_i) mod l], k the innermost index of
 * ix; for a vector that is the DFT, for a higher rank each row receives the same sums.
 *   afft_dft    dft (Forward, rootsOfUnity), idft (Inverse: inverse roots, divided by n), or the
 *               unnormalised inverse (Reverse); roots from Roots.hs:26-51 with the phase reduced in
 *               integers and evaluated in FP64
 *   afft_dftG   caller-supplied roots of the input's shape
 *   afft_dftGS  the single value at `index` (rank entries) into d_out[0]
 * The innermost extent must be below 2^31; out may not alias the input. */
int afft_dft(int mode, int dtype, int rank, const int64_t *extents, const void *d_in, void *d_out,
             void *stream);
int afft_dftG(int dtype, int rank, const int64_t *extents, const void *d_roots, const void *d_in,
              void *d_out, void *stream);
int afft_dftGS(int dtype, int rank, const int64_t *extents, const int64_t *index, const void *d_roots,
               const void *d_in, void *d_out, void *stream);

/* ---- explicit plans --------------------------------------------------------------------- */

/* Build an immutable plan: kernels chosen, twiddles generated on the device in FP64 and rounded
 * to the element type.  `axes` is AFFT_AXES_INNERMOST, AFFT_AXES_ALL or AFFT_AXES_LAST2 (rank
 * >= 2).  `device` is a HIP device ordinal.
 * Concurrency: a plan carries no stream state, and the intermediate arrays of an execution (the
 * work array of a split axis, the auxiliary arrays of a multi-kernel Bluestein transform) are
 * leased per execution from a small per-plan pool, one set per stream in flight (an execution
 * on another stream waits for a set only when four are in flight at once).  So one plan -- a
 * cached one included -- may be executed from several host threads on several streams at once.
 * An execution captured into a graph keeps a set of its own for as long as the plan lives.
 * Replaces the cuFFT handle creation of LLVM/PTX.hs:137-170 and PTX/Plans.hs:40-90. */
int afft_plan_create(afft_plan **plan, int dtype, int rank, const int64_t *extents, int axes,
                     int device);

/* The shared cached plan for a key (what afft_fft* use).  The pointer stays valid until
 * afft_cache_clear; do not destroy it.  Replaces PTX/Plans.hs:66-86 (withPlan). */
int afft_plan_cached(const afft_plan **plan, int dtype, int rank, const int64_t *extents, int axes,
                     int device);

/* Run a plan.  direction: -1 (Forward) or +1 (Reverse), unnormalised, then every output element
 * is multiplied by `scale` inside the last pass (1.0 = the reference plugin contract; 1/N gives
 * the normalised Inverse).  Replaces cuFFT's execC2C/execZ2Z (LLVM/PTX.hs:112-132). */
int afft_exec(const afft_plan *plan, int direction, const void *d_in, void *d_out, double scale,
              void *stream);

int afft_plan_destroy(afft_plan *plan);

/* afft_exec without any device allocation: the intermediate arrays come from a fixed scratch array
 * of AFFT_NOALLOC_SCRATCH_BYTES per GPU (a static device array of the library, held until the
 * process ends) instead of the plan's per-stream pool.  Plans whose intermediates exceed it (see
 * afft_plan_info: device bytes beyond the tables) return AFFT_ERROR_SCRATCH_TOO_SMALL before any
 * launch.  Every call on one GPU shares that GPU's scratch, so calls that overlap on a GPU (two
 * streams or two threads at once) race on it: the caller serialises them.  The plan itself is
 * created beforehand (afft_plan_create / afft_plan_cached allocate its tables).
 * No counterpart in the reference: for callers that cannot allocate device memory during
 * execution (code under graph capture, tight serving loops, harnesses that forbid hipMalloc). */
#define AFFT_NOALLOC_SCRATCH_BYTES (64ll << 20)
int afft_exec_noalloc(const afft_plan *plan, int direction, const void *d_in, void *d_out,
                      double scale, void *stream);

/* Drop every cached plan (and its device memory).  Not safe while another thread is executing a
 * cached plan. */
int afft_cache_clear(void);

/* Number of kernel launches one afft_exec of the plan makes, and the bytes of device memory it
 * needs: its tables (twiddles, Bluestein filters, row tables) plus the intermediate arrays of one
 * execution (each stream executing the plan concurrently leases one such set). */
int afft_plan_info(const afft_plan *plan, int *launches, int64_t *device_bytes);

/* ---- host-only planning (no device needed) ---------------------------------------------- */

/* Describe how a single-workgroup LDS transform of length n would be laid out: the radix of each
 * Stockham pass (at most max_passes written), elements per thread, and the Bluestein padded
 * length (0 when n is 5-smooth).  Returns AFFT_ERROR_UNSUPPORTED_SIZE when n is beyond the
 * single-workgroup limit for the dtype (the multi-pass four-step path then applies). */
int afft_describe(int64_t n, int dtype, int *radices, int max_passes, int *n_passes,
                  int *elems_per_thread, int64_t *bluestein_len);

/* ---- extensions (no counterpart in the reference API) ------------------------------------ */

/* VJP of a front-end op with respect to its input: given the cotangent d_gy of y = op(mode, x),
 * writes d_gx = op^H(d_gy) (the conjugate-transpose transform: Forward <-> Reverse, with the 1/N
 * of Inverse kept).  `op` is 0 = fft (innermost), 1 = fft1D, 2 = fft2D, 3 = fft3D; `extents`
 * has `rank` entries. */
int afft_fft_vjp(int op, int mode, int dtype, int rank, const int64_t *extents, const void *d_gy,
                 void *d_gx, void *stream);

/* Ragged batch: `count` independent 1D transforms of the given lengths, stored back to back
 * (row i starts at element sum_{j<i} lengths[j]).  The lengths that fit one workgroup run in one
 * persistent kernel that drains a device-side work list; every longer length runs as a batched
 * plan over its rows, gathered by offset, on the same stream (latency-bound parts on a forked and
 * rejoined side stream).  In place (d_in == d_out) is allowed.  Extension: the reference has no
 * ragged entry point (FFT.hs:63-84 batches only equal rows). */
int afft_ragged_plan_create(afft_ragged_plan **plan, int dtype, int64_t count,
                            const int64_t *lengths, int device);
int afft_ragged_exec(const afft_ragged_plan *plan, int mode, const void *d_in, void *d_out,
                     void *stream);
/* The same work captured once into a hipGraph per (mode, d_in, d_out) and replayed (kernel nodes
 * only, no memset). */
int afft_ragged_exec_graph(afft_ragged_plan *plan, int mode, const void *d_in, void *d_out,
                           void *stream);
int afft_ragged_plan_destroy(afft_ragged_plan *plan);
/* Total elements of the batch (sum of lengths). */
int afft_ragged_total(const afft_ragged_plan *plan, int64_t *total);

/* Spectral convolution over a batch of h x w images:
 *   y_b = Inverse(fft2D(x_b) * H)  (elementwise product with the h x w filter H).
 * Fused: row pass, then one kernel that does the column transform, the filter product and the
 * inverse column transform in LDS, then the inverse row pass with the 1/(h w) scale. */
int afft_spectral_conv_fwd(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                           const void *d_H, void *d_y, void *stream);

/* Its VJP: given d_gy = dL/dy (conjugate-Wirtinger convention, as torch), writes
 *   d_gx = (1/hw) F^H (conj(H) * F gy)        and
 *   d_gH = (1/hw) sum_b conj(F x_b) * F gy_b
 * recomputing F x_b instead of storing it.  Either output pointer may be NULL to skip it. */
int afft_spectral_conv_vjp(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                           const void *d_H, const void *d_gy, void *d_gx, void *d_gH,
                           void *stream);

/* The forward that also stores F x_b, the unnormalised 2D spectrum of every image (B x h x w), in
 * d_spec: the array the VJP otherwise recomputes, written from the column kernel's registers (one
 * more array written).  d_spec must not be d_y or d_H. */
int afft_spectral_conv_fwd_save(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                const void *d_H, void *d_y, void *d_spec, void *stream);

/* The VJP from a stored spectrum d_spec = F x (as afft_spectral_conv_fwd_save writes it): no row
 * pass over x and one column transform less per image.  d_spec must not be d_gx or d_gH. */
int afft_spectral_conv_vjp_saved(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_spec,
                                 const void *d_H, const void *d_gy, void *d_gx, void *d_gH,
                                 void *stream);

/* The VJP of a squared-error loss L = (alpha / 2) sum_b |y_b - t_b|^2 through the convolution,
 * i.e. with dL/dy = alpha (d_y - d_t), which the first row pass forms as it loads y and t: the
 * upstream gradient is never stored (alpha = 2 for sum |y - t|^2).  d_x is x, or with
 * x_is_spectrum = 1 the stored spectrum F x (as for afft_spectral_conv_vjp_saved).  Outputs as
 * afft_spectral_conv_vjp.  No counterpart in the reference API (an extension of the adjoint
 * kernels for training loops, like the two entry points above). */
int afft_spectral_conv_vjp_residual(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                    int x_is_spectrum, const void *d_H, const void *d_y,
                                    const void *d_t, double alpha, void *d_gx, void *d_gH,
                                    void *stream);

/* Autotuning: time the candidate layouts of every decision the planner makes for a shape (the row
 * kernel variant, elements per thread and columns per workgroup of strided tiles -- which fix the
 * radices and their order --, the lanes per column of the register kernel) on the device, one
 * decision after the other, and remember the fastest; plans built afterwards use them (the plan
 * cache is cleared).  afft_ragged_autotune does the same for every length class of a ragged batch,
 * each timed alone (for a class of the persistent kernel its compiled layout against the
 * pass-table layouts of each cap on elements per thread, timed on at least 2^23 elements; the
 * layouts of the long classes' plans), then for the whole batch which long classes run on the
 * side stream (keyed by a hash of the batch's length classes).  afft_tune_save/afft_tune_load
 * persist the choices as a text file keyed by (gfx arch, HIP runtime version, decision kind,
 * dtype, length); stale or malformed lines are ignored. */
int afft_autotune(int dtype, int rank, const int64_t *extents, int axes, int device);
int afft_ragged_autotune(int dtype, int64_t count, const int64_t *lengths, int device);
int afft_tune_save(const char *path);
int afft_tune_load(const char *path);

#ifdef __cplusplus
}
#endif

#endif /* AFFT_H */
