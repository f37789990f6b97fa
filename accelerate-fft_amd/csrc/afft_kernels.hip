// afft_kernels.hip -- gfx950 kernels of the FFT engine.
//
//   afft_stockham_<f32|f64>_e<E>   batched single-workgroup LDS Stockham transform (rows,
//                                  strided axes, and both passes of the four-step scheme)
//   afft_bluestein_<f32|f64>_e<E>  fused chirp-z transform for lengths that are not 5-smooth
//   afft_gen_*                     FP64 twiddle / chirp generation, rounded to the element type
//
// The path is HBM-bound (about 0.3 log2 N flop per byte), so nothing here uses MFMA: the design
// goal is one HBM read and one HBM write per element per launch, with the passes of a transform
// exchanged through LDS.
#include <hip/hip_runtime.h>

#include "afft_device.hpp"
#include "afft_internal.hpp"

namespace afft {

template <int E> struct ElemsTag {};

// thread -> (transform slot c in the workgroup, thread t within the transform)
__device__ __forceinline__ void split_tid(int tid, int C, int T, int lane_col, int &c, int &t)
{
    if (lane_col) {
        c = tid % C;
        t = tid / C;
    } else {
        t = tid % T;
        c = tid / T;
    }
}

template <typename V>
__device__ __forceinline__ V *lds_slot(V *lds, int c, int C, int n, int interleave, int &es)
{
    if (interleave) {
        es = C;
        return lds + c;
    }
    es = 1;
    return lds + (int64_t)c * lds_padded_len(n);
}

// ----------------------------------------------------------------------------------------------
// Batched Stockham transform.  N = 0: the generic transform of any supported n, driven by the pass
// table; N > 0: n = N with compile-time radices Rs (CtStockham, the row kernels' passes; used for
// the interleaved tiles of strided axes), twiddle bases loaded once from the table W_N^m.
// TWG: the four-step twiddles on store come from the precomputed grid post_tab, loaded with the
// data so that they are in flight during the transform instead of in front of every store (the
// middle pass of the split six-step); otherwise from the two-level table post_hi / post_lo.

template <typename Real, int E, int N, bool TWG, int... Rs>
__global__ void __launch_bounds__(max_block_threads(sizeof(Real) == 8, E)) afft_stockham(const FftArgs a)
{
    static_assert(!TWG || N != 0, "the twiddle grid goes with the compile-time tiles");
    using V = typename cplx_of<Real>::type;
    extern __shared__ __align__(16) unsigned char afft_smem[];
    V *lds = reinterpret_cast<V *>(afft_smem);

    const int n = N ? N : a.pt.n;
    const int T = n / E;
    int c, t;
    split_tid(threadIdx.x, a.C, T, a.lane_col, c, t);
    CtStockham<V, N ? N : E, E, Rs...> ct;
    if constexpr (N != 0) ct.load(reinterpret_cast<const V *>(a.pw), t);
    const int64_t tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t j = tile * a.C + c;
    const bool active = (c < a.C) && (j < a.ntrans);

    V reg[E];
    if (active) {
        const V *in = reinterpret_cast<const V *>(a.in);
        const int64_t base = addr_base(a.ain, j);
        const bool nt = a.nt_in;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            V v = c_load(in + base + (int64_t)(t + i * T) * a.ain.es, nt);
            reg[i] = a.conj_in ? c_conj(v) : v;
        }
        if (a.sg_in.on) { // apart from the loads, which stay one batch of independent instructions
#pragma unroll
            for (int i = 0; i < E; ++i) reg[i] = sign_apply(a.sg_in, base + (int64_t)(t + i * T) * a.ain.es, reg[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < E; ++i) reg[i] = c_make<V>(0, 0);
    }
    V twv[TWG ? E : 1];
    if constexpr (TWG) {
        if (active) {
            const V *tab = reinterpret_cast<const V *>(a.post_tab) + addr_level(a.ain, j, a.post_level);
#pragma unroll
            for (int i = 0; i < E; ++i) twv[i] = tab[(int64_t)(t + i * T) * a.post_ld];
        }
    }

    int es;
    V *slot = lds_slot(lds, c < a.C ? c : 0, a.C, n, a.lds_interleave, es);
    if constexpr (N != 0) ct.run(reg, slot, es, t);
    else fft_core<E>(reg, slot, es, a.pt, reinterpret_cast<const V *>(a.tw), t, T);

    if (active) {
        V *out = reinterpret_cast<V *>(a.out);
        const int64_t base = addr_base(a.aout, j);
        const Real sc = (Real)a.scale;
        const int64_t jl = a.post_L ? addr_level(a.ain, j, a.post_level) : 0;
        if (a.sg_out.on) { // apart from the stores, which stay one batch
#pragma unroll
            for (int i = 0; i < E; ++i) reg[i] = sign_apply(a.sg_out, base + (int64_t)(t + i * T) * a.aout.es, reg[i]);
        }
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int k = t + i * T;
            V v = reg[i];
            if constexpr (TWG) {
                v = c_mul(v, twv[TWG ? i : 0]);
            } else if (a.post_L) {
                // jl < n2 and k < n1, so jl * k < n1 * n2 = post_L: no reduction needed
                const int64_t e = jl * (int64_t)k;
                v = c_mul(v, power_lookup(reinterpret_cast<const V *>(a.post_hi),
                                          reinterpret_cast<const V *>(a.post_lo), e, a.post_shift));
            }
            if (a.conj_out) v = c_conj(v);
            c_store(out + base + (int64_t)k * a.aout.es, c_scale(v, sc), a.nt_out);
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Fused Bluestein: X[k] = c[k] * IFFT_M( FFT_M(x c) * FFT_M(b)/M )[k],  c[m] = exp(-i pi m^2/n).
// The inverse FFT is conj(FFT(conj(.))) so both transforms use the forward passes.  MC = 0: the
// transforms of M driven by the pass table; MC > 0: M = MC with compile-time radices (CtStockham).

template <typename Real, int E, int MC, int... Rs>
__global__ void __launch_bounds__(max_block_threads(sizeof(Real) == 8, E)) afft_bluestein(const BluesteinArgs a)
{
    using V = typename cplx_of<Real>::type;
    extern __shared__ __align__(16) unsigned char afft_smem[];
    V *lds = reinterpret_cast<V *>(afft_smem);

    const int M = MC ? MC : a.pt.n;
    const int T = M / E;
    int c, t;
    split_tid(threadIdx.x, a.C, T, a.lane_col, c, t);
    CtStockham<V, MC ? MC : E, E, Rs...> ct;
    if constexpr (MC != 0) ct.load(reinterpret_cast<const V *>(a.pw), t);
    const int64_t tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t j = tile * a.C + c;
    const bool active = (c < a.C) && (j < a.ntrans);
    const V *chirp = reinterpret_cast<const V *>(a.chirp);
    const V *filt = reinterpret_cast<const V *>(a.filt);
    const V *tw = reinterpret_cast<const V *>(a.tw);

    V reg[E];
    if (active) {
        const V *in = reinterpret_cast<const V *>(a.in);
        const int64_t base = addr_base(a.ain, j);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int m = t + i * T;
            V v = c_make<V>(0, 0);
            if (m < a.n) {
                const int64_t off = base + (int64_t)m * a.ain.es;
                v = sign_apply(a.sg_in, off, c_load(in + off, (bool)a.nt_in));
                if (a.conj_in) v = c_conj(v);
                v = c_mul(v, chirp[m]);
            }
            reg[i] = v;
        }
    } else {
#pragma unroll
        for (int i = 0; i < E; ++i) reg[i] = c_make<V>(0, 0);
    }

    // the filter is read with the data when the registers allow it (the 512-thread layout of
    // 10240 keeps two waves per SIMD with it), so that its loads are in flight during the first
    // transform instead of in front of the second
    constexpr bool kEarlyFilt = MC != 0 && E <= 20;
    V fv[kEarlyFilt ? E : 1];
    if constexpr (kEarlyFilt) {
#pragma unroll
        for (int i = 0; i < E; ++i) fv[i] = filt[t + i * T];
    }
    int es;
    V *slot = lds_slot(lds, c < a.C ? c : 0, a.C, M, a.lds_interleave, es);
    if constexpr (MC != 0) {
        // padded positions in a private slot; the thread index passes through an empty asm before
        // each transform so that the LDS addresses are not all kept live across both
        int tl = t;
        __asm__ volatile("" : "+v"(tl));
        if (a.lds_interleave) ct.run(reg, slot, es, tl);
        else ct.template run<true>(reg, slot, es, tl);
    } else {
        fft_core<E>(reg, slot, es, a.pt, tw, t, T);
    }
    if constexpr (kEarlyFilt) {
#pragma unroll
        for (int i = 0; i < E; ++i) reg[i] = c_conj(c_mul(reg[i], fv[kEarlyFilt ? i : 0]));
    } else {
#pragma unroll
        for (int i = 0; i < E; ++i) reg[i] = c_conj(c_mul(reg[i], filt[t + i * T]));
    }
    if constexpr (MC != 0) {
        int tl = t;
        __asm__ volatile("" : "+v"(tl));
        if (a.lds_interleave) ct.run(reg, slot, es, tl);
        else ct.template run<true>(reg, slot, es, tl);
    } else {
        fft_core<E>(reg, slot, es, a.pt, tw, t, T);
    }

    if (active) {
        V *out = reinterpret_cast<V *>(a.out);
        const int64_t base = addr_base(a.aout, j);
        const Real sc = (Real)a.scale;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int m = t + i * T;
            if (m < a.n) {
                V v = c_mul(c_conj(reg[i]), chirp[m]);
                if (a.conj_out) v = c_conj(v);
                const int64_t off = base + (int64_t)m * a.aout.es;
                c_store(out + off, sign_apply(a.sg_out, off, c_scale(v, sc)), (bool)a.nt_out);
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Pointwise stages of the multi-kernel Bluestein path.

template <typename Real>
__global__ void __launch_bounds__(256) afft_chirp(const ChirpArgs a)
{
    using V = typename cplx_of<Real>::type;
    const V *chirp = reinterpret_cast<const V *>(a.chirp);
    const int64_t per = a.mode == 2 ? a.n : a.M;
    const int64_t total = a.ntrans * per;
    for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
         idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = idx / per;
        const int64_t m = idx - j * per;
        if (a.mode == 0) {
            const V *x = reinterpret_cast<const V *>(a.in);
            V *aux = reinterpret_cast<V *>(a.out);
            V v = c_make<V>(0, 0);
            if (m < a.n) {
                const int64_t off = addr_base(a.a, j) + m * a.a.es;
                v = sign_apply(a.sg, off, c_load(x + off, (bool)a.nt_in));
                if (a.conj_in) v = c_conj(v);
                v = c_mul(v, chirp[m]);
            }
            aux[j * a.M + m] = v;
        } else if (a.mode == 1) {
            V *aux = reinterpret_cast<V *>(a.out);
            const V *filt = reinterpret_cast<const V *>(a.filt);
            aux[j * a.M + m] = c_conj(c_mul(aux[j * a.M + m], filt[m]));
        } else {
            const V *aux = reinterpret_cast<const V *>(a.in);
            V *y = reinterpret_cast<V *>(a.out);
            V v = c_mul(c_conj(aux[j * a.M + m]), chirp[m]);
            if (a.conj_out) v = c_conj(v);
            const int64_t off = addr_base(a.a, j) + m * a.a.es;
            c_store(y + off, sign_apply(a.sg, off, c_scale(v, (Real)a.scale)), (bool)a.nt_out);
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Twiddle / chirp generation in FP64.

template <typename Real>
__global__ void afft_gen_pass_twiddles(typename cplx_of<Real>::type *out, const PassTable pt)
{
    using V = typename cplx_of<Real>::type;
    int Ns = 1;
    for (int p = 0; p < pt.npass; ++p) {
        const int R = pt.radix[p];
        const int L = Ns * R;
        for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < L; idx += gridDim.x * blockDim.x) {
            const int k = idx / R, r = idx % R;
            const int e = r * k; // < L, exact
            double s, co;
            sincospi(-2.0 * (double)e / (double)L, &s, &co);
            out[pt.tw_off[p] + idx] = c_make<V>((Real)co, (Real)s);
        }
        Ns = L;
    }
}

template <typename Real>
__global__ void afft_gen_power_table(typename cplx_of<Real>::type *out, int64_t L, int64_t step,
                                     int64_t count)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        // exponent reduced exactly in integers; i*step < 2^62 for every table built here
        const int64_t e = (i * step) % L;
        double s, co;
        sincospi(-2.0 * (double)e / (double)L, &s, &co);
        out[i] = c_make<V>((Real)co, (Real)s);
    }
}

template <typename Real>
__global__ void afft_gen_twiddle_grid(typename cplx_of<Real>::type *out, int64_t L, int64_t rows, int64_t cols)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < rows * cols;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = i / cols, j = i - q * cols;
        const int64_t e = (q * j) % L;
        double s, co;
        sincospi(-2.0 * (double)e / (double)L, &s, &co);
        out[i] = c_make<V>((Real)co, (Real)s);
    }
}

template <typename Real>
__global__ void afft_gen_chirp(typename cplx_of<Real>::type *out, int64_t n)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < n;
         m += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = (m * m) % (2 * n); // exp(-i pi m^2/n) has period 2n in m^2
        double s, co;
        sincospi(-(double)e / (double)n, &s, &co);
        out[m] = c_make<V>((Real)co, (Real)s);
    }
}

__global__ void afft_gen_bluestein_kernel(double2 *out, int64_t n, int64_t M)
{
    for (int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; m < M;
         m += (int64_t)gridDim.x * blockDim.x) {
        int64_t src = -1;
        if (m < n) src = m;
        else if (M - m < n) src = M - m;
        double2 v = make_double2(0.0, 0.0);
        if (src >= 0) {
            const int64_t e = (src * src) % (2 * n);
            double s, co;
            sincospi(-(double)e / (double)n, &s, &co);
            v = make_double2(co, -s); // conj(chirp)
        }
        out[m] = v;
    }
}

template <typename Real>
__global__ void afft_convert_scale(typename cplx_of<Real>::type *out, const double2 *in,
                                   int64_t count, double scale)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double2 v = in[i];
        out[i] = c_make<V>((Real)(v.x * scale), (Real)(v.y * scale));
    }
}

template <typename Real>
__global__ void afft_pointwise_mul(typename cplx_of<Real>::type *y, const typename cplx_of<Real>::type *x,
                                   const typename cplx_of<Real>::type *h, int64_t count, int64_t hcount,
                                   int conj_h, Real scale)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        V hv = h[i % hcount];
        if (conj_h) hv = c_conj(hv);
        y[i] = c_scale(c_mul(x[i], hv), scale);
    }
}

// ----------------------------------------------------------------------------------------------
// Host launchers.

const int kElemsPerThread[] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16, 18, 20, 24, 25, 27, 30, 32, 40, 48};
const int kNumElemsPerThread = sizeof(kElemsPerThread) / sizeof(kElemsPerThread[0]);

#define AFFT_FOR_EACH_E(X)                                                                       \
    X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16) X(18) X(20) X(24) X(25) X(27)  \
    X(30) X(32) X(40) X(48)

template <typename Kern, typename Args>
static hipError_t launch_one(Kern k, const Args &a, int block, int64_t grid, size_t lds, hipStream_t s)
{
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(k),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(block), lds, s, a);
    return hipGetLastError();
}

template <typename Real>
static hipError_t launch_stockham_t(const FftArgs &a, int block, int64_t grid, size_t lds, hipStream_t s)
{
    // compile-time transforms for the interleaved tiles of strided axes (the radices the layout
    // chooser picks for these lengths)
    if (a.pw && a.lds_interleave && a.E == 16) {
        if constexpr (sizeof(Real) == 4) {
            // the twiddle-grid forms in Complex Float only: in FP64 its registers would cost the
            // strided tiles a wave per SIMD
            if (a.post_tab) {
                if (a.pt.n == 128) return launch_one(afft_stockham<Real, 16, 128, true, 16, 8>, a, block, grid, lds, s);
                if (a.pt.n == 256) return launch_one(afft_stockham<Real, 16, 256, true, 16, 16>, a, block, grid, lds, s);
                if (a.pt.n == 512) return launch_one(afft_stockham<Real, 16, 512, true, 16, 16, 2>, a, block, grid, lds, s);
                if (a.pt.n == 1024) return launch_one(afft_stockham<Real, 16, 1024, true, 16, 16, 4>, a, block, grid, lds, s);
            }
            if (a.pt.n == 128) return launch_one(afft_stockham<Real, 16, 128, false, 16, 8>, a, block, grid, lds, s);
            if (a.pt.n == 512) return launch_one(afft_stockham<Real, 16, 512, false, 16, 16, 2>, a, block, grid, lds, s);
            if (a.pt.n == 1024) return launch_one(afft_stockham<Real, 16, 1024, false, 16, 16, 4>, a, block, grid, lds, s);
        }
        if (a.pt.n == 256) return launch_one(afft_stockham<Real, 16, 256, false, 16, 16>, a, block, grid, lds, s);
    }
    switch (a.E) {
#define X(EV)                                                                                    \
    case EV: return launch_one(afft_stockham<Real, EV, 0, false>, a, block, grid, lds, s);
        AFFT_FOR_EACH_E(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

template <typename Real>
static hipError_t launch_bluestein_t(const BluesteinArgs &a, int block, int64_t grid, size_t lds,
                                     hipStream_t s)
{
    if constexpr (sizeof(Real) == 4) {
        // compile-time transforms for the padded lengths of two common primes (257 -> 640,
        // 4099 -> 10240): the radices the layout chooser picks, and for 10240 also the
        // 512-thread layout of contiguous rows
        if (a.pw) {
            if (a.pt.n == 640 && a.E == 40) return launch_one(afft_bluestein<Real, 40, 640, 8, 8, 5, 2>, a, block, grid, lds, s);
            if (a.pt.n == 10240 && a.E == 40) return launch_one(afft_bluestein<Real, 40, 10240, 8, 8, 8, 5, 4>, a, block, grid, lds, s);
            if (a.pt.n == 10240 && a.E == 20) return launch_one(afft_bluestein<Real, 20, 10240, 2, 4, 4, 4, 4, 4, 5>, a, block, grid, lds, s);
        }
    }
    switch (a.E) {
#define X(EV)                                                                                    \
    case EV: return launch_one(afft_bluestein<Real, EV, 0>, a, block, grid, lds, s);
        AFFT_FOR_EACH_E(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stockham(int dtype, const FftArgs &a, int block, int64_t grid, size_t lds, hipStream_t s)
{
    return dtype == 0 ? launch_stockham_t<float>(a, block, grid, lds, s)
                      : launch_stockham_t<double>(a, block, grid, lds, s);
}

hipError_t launch_bluestein(int dtype, const BluesteinArgs &a, int block, int64_t grid, size_t lds,
                            hipStream_t s)
{
    return dtype == 0 ? launch_bluestein_t<float>(a, block, grid, lds, s)
                      : launch_bluestein_t<double>(a, block, grid, lds, s);
}

static int64_t gs_grid(int64_t count)
{
    int64_t g = (count + 255) / 256;
    if (g < 1) g = 1;
    if (g > 4096) g = 4096;
    return g;
}

hipError_t gen_pass_twiddles(int dtype, void *out, const PassTable &pt, hipStream_t s)
{
    const int64_t g = gs_grid(pt.n);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_gen_pass_twiddles<float>, dim3(g), dim3(256), 0, s, (float2 *)out, pt);
    else
        hipLaunchKernelGGL(afft_gen_pass_twiddles<double>, dim3(g), dim3(256), 0, s, (double2 *)out, pt);
    return hipGetLastError();
}

hipError_t gen_power_table(int dtype, void *out, int64_t L, int64_t step, int64_t count, hipStream_t s)
{
    const int64_t g = gs_grid(count);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_gen_power_table<float>, dim3(g), dim3(256), 0, s, (float2 *)out, L, step, count);
    else
        hipLaunchKernelGGL(afft_gen_power_table<double>, dim3(g), dim3(256), 0, s, (double2 *)out, L, step, count);
    return hipGetLastError();
}

hipError_t gen_twiddle_grid(int dtype, void *out, int64_t L, int64_t rows, int64_t cols, hipStream_t s)
{
    const int64_t g = gs_grid(rows * cols);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_gen_twiddle_grid<float>, dim3(g), dim3(256), 0, s, (float2 *)out, L, rows, cols);
    else
        hipLaunchKernelGGL(afft_gen_twiddle_grid<double>, dim3(g), dim3(256), 0, s, (double2 *)out, L, rows, cols);
    return hipGetLastError();
}

hipError_t gen_chirp(int dtype, void *out, int64_t n, hipStream_t s)
{
    const int64_t g = gs_grid(n);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_gen_chirp<float>, dim3(g), dim3(256), 0, s, (float2 *)out, n);
    else
        hipLaunchKernelGGL(afft_gen_chirp<double>, dim3(g), dim3(256), 0, s, (double2 *)out, n);
    return hipGetLastError();
}

hipError_t gen_bluestein_kernel_f64(void *out, int64_t n, int64_t M, hipStream_t s)
{
    hipLaunchKernelGGL(afft_gen_bluestein_kernel, dim3(gs_grid(M)), dim3(256), 0, s, (double2 *)out, n, M);
    return hipGetLastError();
}

hipError_t convert_scale_from_f64(int dtype, void *out, const void *in, int64_t count, double scale,
                                  hipStream_t s)
{
    const int64_t g = gs_grid(count);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_convert_scale<float>, dim3(g), dim3(256), 0, s, (float2 *)out,
                           (const double2 *)in, count, scale);
    else
        hipLaunchKernelGGL(afft_convert_scale<double>, dim3(g), dim3(256), 0, s, (double2 *)out,
                           (const double2 *)in, count, scale);
    return hipGetLastError();
}

hipError_t launch_chirp(int dtype, const ChirpArgs &a, hipStream_t s)
{
    const int64_t per = a.mode == 2 ? a.n : a.M;
    int64_t g = (a.ntrans * per + 255) / 256;
    if (g > 65536) g = 65536;
    if (g < 1) return hipSuccess;
    if (dtype == 0)
        hipLaunchKernelGGL(afft_chirp<float>, dim3((unsigned)g), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(afft_chirp<double>, dim3((unsigned)g), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t pointwise_mul(int dtype, void *y, const void *x, const void *h, int64_t count,
                         int64_t hcount, int conj_h, double scale, hipStream_t s)
{
    const int64_t g = gs_grid(count);
    if (dtype == 0)
        hipLaunchKernelGGL(afft_pointwise_mul<float>, dim3(g), dim3(256), 0, s, (float2 *)y,
                           (const float2 *)x, (const float2 *)h, count, hcount, conj_h, (float)scale);
    else
        hipLaunchKernelGGL(afft_pointwise_mul<double>, dim3(g), dim3(256), 0, s, (double2 *)y,
                           (const double2 *)x, (const double2 *)h, count, hcount, conj_h, scale);
    return hipGetLastError();
}

} // namespace afft
