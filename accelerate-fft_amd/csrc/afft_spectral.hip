// afft_spectral.hip -- fused spectral convolution over a batch of h x w images and its VJP
// (afft_spectral_conv_fwd / afft_spectral_conv_vjp; no counterpart in the reference API).
//
//   forward   y_b = (1/hw) F^H (H * F x_b)       F = unnormalised 2D DFT, F^H = Reverse mode
//   VJP       gx_b = (1/hw) F^H (conj(H) * F gy_b)
//             gH   = (1/hw) sum_b conj(F x_b) * F gy_b
//
// F x_b is either recomputed by the VJP (afft_spectral_conv_vjp: no memory beyond the call) or
// stored by the forward call, which has it in registers between its column transform and the
// filter product (afft_spectral_conv_fwd_save: one more array written), and read back by the
// VJP (afft_spectral_conv_vjp_saved: no row pass over x, one column transform less per image).
//
// Launch structure (forward): row pass over all B*h rows (the persistent row kernel), then ONE
// column kernel per column tile that does the column FFT, the product with H and the inverse
// column FFT in LDS (afft_colconv), then the inverse row pass carrying the 1/hw scale.  Three
// array passes instead of the six of the per-call composition (2D forward, product, 2D
// inverse, scale).  The VJP column kernel (afft_colconv_vjp) transforms the columns of both
// row-transformed arrays, accumulates conj(X) G for its group of images in registers (a
// deterministic reduction, no atomics) and writes the back-transformed conj(H) G.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "../../include/afft.h"
#include "afft_device.hpp"
#include "afft_internal.hpp"
#include "afft_plan.hpp"

namespace afft {

struct ColConvArgs {
    void *data;          // B x h x w, row-transformed; column-transformed result written back
    const void *other;   // VJP: the row-transformed x, or F x itself when other_is_spec (B x h x w)
    void *spec;          // forward: where F x is stored (B x h x w), or null
    const void *H;       // h x w filter
    void *partial;       // VJP: groups x h x w partial sums of conj(X) G
    const void *tw;
    PassTable pt;        // length h
    const void *pw;      // W_h^m, m in [0, h): the twiddle bases of the compile-time transforms
    int64_t B, h, w;
    int C;               // columns per tile
    int groups;          // VJP: images are split into `groups` contiguous groups
    int other_is_spec;   // VJP: `other` holds F x (column transform already done)
};

// Workgroup -> (outer index, column tile).  The dispatcher deals workgroup b to XCD b mod 8; each
// XCD takes a fixed eighth of the column tiles for every image (or image group), so it reads only
// its eighth of the filter H, which then stays in that XCD's L2 from image to image instead of
// being fetched again for every image (measured: forward 171 -> 164 us on 64 x 512^2; the VJP,
// bound elsewhere, unchanged).  Falls back to the XCD-contiguous tile order when the
// column tiles do not split evenly.
__device__ __forceinline__ void colconv_tile(int64_t tiles, int64_t outer, int64_t &o, int64_t &ct)
{
    constexpr int kXcd = 8;
    const int64_t b = blockIdx.x;
    if (tiles % kXcd == 0 && (int64_t)gridDim.x == tiles * outer) {
        const int64_t per = tiles / kXcd;
        const int64_t k = b / kXcd;
        o = k / per;
        ct = (b % kXcd) * per + k % per;
        return;
    }
    const int64_t tile = xcd_tile(b, gridDim.x);
    o = tile / tiles;
    ct = tile % tiles;
}

// The column transform of a tile: h = N with compile-time radices Rs (the Stockham passes of the
// row kernels, twiddle bases loaded once per workgroup), or, for N = 0, the generic transform
// of any 5-smooth h driven by the pass table.
template <typename V, int E, int N, int... Rs> struct ColFft {
    CtStockham<V, N, E, Rs...> ct;
    __device__ __forceinline__ void init(const ColConvArgs &a, int t) { ct.load(reinterpret_cast<const V *>(a.pw), t); }
    // the thread index passes through an empty asm, so each transform recomputes its LDS
    // addresses instead of the compiler keeping all of them live across both transforms
    __device__ __forceinline__ void run(V (&reg)[E], V *base, const ColConvArgs &a, int t) const
    {
        int tl = t;
        __asm__ volatile("" : "+v"(tl));
        ct.run(reg, base, a.C, tl);
    }
};
template <typename V, int E> struct ColFft<V, E, 0> {
    __device__ __forceinline__ void init(const ColConvArgs &, int) {}
    __device__ __forceinline__ void run(V (&reg)[E], V *base, const ColConvArgs &a, int t) const
    {
        fft_core<E>(reg, base, a.C, a.pt, reinterpret_cast<const V *>(a.tw), t, (int)(a.h / E));
    }
};

template <typename Real, int E, int N, int... Rs>
__global__ void __launch_bounds__(max_block_threads(sizeof(Real) == 8, E)) afft_colconv(const ColConvArgs a)
{
    using V = typename cplx_of<Real>::type;
    extern __shared__ __align__(16) unsigned char afft_smem[];
    V *lds = reinterpret_cast<V *>(afft_smem);
    const int T = N ? N / E : (int)(a.h / E), C = a.C;
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int64_t tiles = (a.w + C - 1) / C;
    int64_t b, ct;
    colconv_tile(tiles, a.B, b, ct);
    const int64_t j = ct * C + c;
    const bool active = j < a.w && b < a.B;
    V *d = reinterpret_cast<V *>(a.data) + b * a.h * a.w + j;
    const V *Hc = reinterpret_cast<const V *>(a.H) + j;
    ColFft<V, E, N, Rs...> fft;
    fft.init(a, t);
    V reg[E];
#pragma unroll
    for (int i = 0; i < E; ++i) reg[i] = active ? d[(int64_t)(t + i * T) * a.w] : c_make<V>(0, 0);
    fft.run(reg, lds + c, a, t);
    if (a.spec && active) { // F x for a later VJP
        V *sp = reinterpret_cast<V *>(a.spec) + b * a.h * a.w + j;
#pragma unroll
        for (int i = 0; i < E; ++i) sp[(int64_t)(t + i * T) * a.w] = reg[i];
    }
    // product with H, then the inverse column transform as conj(F(conj(.)))
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const V h = active ? Hc[(int64_t)(t + i * T) * a.w] : c_make<V>(0, 0);
        reg[i] = c_conj(c_mul(reg[i], h));
    }
    fft.run(reg, lds + c, a, t);
    if (active) {
#pragma unroll
        for (int i = 0; i < E; ++i) d[(int64_t)(t + i * T) * a.w] = c_conj(reg[i]);
    }
}

// Holds three transforms' worth of registers per thread (x, gy and the running sum): compiled
// for workgroups of at most vjp_block threads so the register budget covers them (512 in FP32:
// 16 columns per tile, whole-line segments, one workgroup of 8 waves per CU).
constexpr int vjp_block(bool f64) { return f64 ? 256 : 512; }

template <typename Real, int E, int N, int... Rs>
__global__ void __launch_bounds__(vjp_block(sizeof(Real) == 8)) __attribute__((amdgpu_waves_per_eu(sizeof(Real) == 4 ? 2 : 1, 8))) afft_colconv_vjp(const ColConvArgs a)
{
    using V = typename cplx_of<Real>::type;
    extern __shared__ __align__(16) unsigned char afft_smem[];
    V *lds = reinterpret_cast<V *>(afft_smem);
    const int T = N ? N / E : (int)(a.h / E), C = a.C;
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const int64_t tiles = (a.w + C - 1) / C;
    int64_t g, ct;
    colconv_tile(tiles, a.groups, g, ct);
    const int64_t j = ct * C + c;
    const bool active = j < a.w;
    const int64_t per = (a.B + a.groups - 1) / a.groups;
    const int64_t b0 = g * per, b1 = (b0 + per < a.B) ? b0 + per : a.B;
    const V *Hc = reinterpret_cast<const V *>(a.H) + j;
    ColFft<V, E, N, Rs...> fft;
    fft.init(a, t);
    V acc[E];
#pragma unroll
    for (int i = 0; i < E; ++i) acc[i] = c_make<V>(0, 0);
    for (int64_t b = b0; b < b1; ++b) {
        // the thread index passes through an empty asm once per image, so the LDS and element
        // addresses are recomputed here instead of all being hoisted out of the loop, where they
        // would hold most of the register file and leave one wave per SIMD
        int tl = t;
        __asm__ volatile("" : "+v"(tl));
        const V *x = reinterpret_cast<const V *>(a.other) + b * a.h * a.w + j;
        V *gy = reinterpret_cast<V *>(a.data) + b * a.h * a.w + j;
        V X[E], G[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            X[i] = active ? x[(int64_t)(tl + i * T) * a.w] : c_make<V>(0, 0);
            G[i] = active ? gy[(int64_t)(tl + i * T) * a.w] : c_make<V>(0, 0);
        }
        if (!a.other_is_spec) fft.run(X, lds + c, a, tl); // uniform branch
        fft.run(G, lds + c, a, tl);
#pragma unroll
        for (int i = 0; i < E; ++i) {
            acc[i] = c_add(acc[i], c_mulc(G[i], X[i])); // conj(X) G
            const V h = active ? Hc[(int64_t)(tl + i * T) * a.w] : c_make<V>(0, 0);
            G[i] = c_conj(c_mulc(G[i], h));             // conj(conj(H) G)
        }
        fft.run(G, lds + c, a, tl);
        if (active) {
#pragma unroll
            for (int i = 0; i < E; ++i) gy[(int64_t)(tl + i * T) * a.w] = c_conj(G[i]);
        }
    }
    if (active) {
        V *p = reinterpret_cast<V *>(a.partial) + g * a.h * a.w + j;
#pragma unroll
        for (int i = 0; i < E; ++i) p[(int64_t)(t + i * T) * a.w] = acc[i];
    }
}

// gH[k] = scale * sum_g partial[g][k]
template <typename Real>
__global__ void afft_reduce_groups(typename cplx_of<Real>::type *out, const typename cplx_of<Real>::type *partial,
                                   int groups, int64_t n, Real scale)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        V s = c_make<V>(0, 0);
        for (int g = 0; g < groups; ++g) s = c_add(s, partial[g * n + k]);
        out[k] = c_scale(s, scale);
    }
}

// out[k] = alpha (y[k] - t[k]): the residual of a squared-error loss, for the paths that need
// the upstream gradient stored
template <typename Real>
__global__ void afft_residual(typename cplx_of<Real>::type *out, const typename cplx_of<Real>::type *y,
                              const typename cplx_of<Real>::type *t, Real alpha, int64_t n)
{
    using V = typename cplx_of<Real>::type;
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x) {
        const V a = y[k], b = t[k];
        out[k] = c_make<V>(alpha * (a.x - b.x), alpha * (a.y - b.y));
    }
}

namespace {

hipError_t launch_residual(int dtype, void *out, const void *y, const void *t, double alpha, int64_t n, hipStream_t s)
{
    const int64_t g = std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256));
    if (dtype == AFFT_C64)
        hipLaunchKernelGGL(afft_residual<float>, dim3((unsigned)g), dim3(256), 0, s, (float2 *)out,
                           (const float2 *)y, (const float2 *)t, (float)alpha, n);
    else
        hipLaunchKernelGGL(afft_residual<double>, dim3((unsigned)g), dim3(256), 0, s, (double2 *)out,
                           (const double2 *)y, (const double2 *)t, alpha, n);
    return hipGetLastError();
}

// The forward row transforms of alpha (y - t) into out: the row plan's single row launch in its
// residual-load form, which reads y and t and never stores the difference; else the difference
// stored into out, then transformed in place.
int residual_rows(const Plan &rp, int dtype, const void *y, const void *t, double alpha, void *out, hipStream_t s)
{
    if (rp.launches.size() == 1 && rp.launches[0].kind == Launch::kRows &&
        row_block(dtype, rp.launches[0].row_n, kRowResidual) == row_block(dtype, rp.launches[0].row_n, rp.launches[0].row_variant) &&
        row_block(dtype, rp.launches[0].row_n, kRowResidual) > 0) {
        const Launch &l = rp.launches[0];
        RowArgs a = l.ra;
        a.in = y;
        a.in2 = t;
        a.alpha = alpha;
        a.out = out;
        a.conj_in = a.conj_out = 0;
        a.scale = 1.0;
        a.nt_in = 1;  // y and t are read once
        a.nt_out = 0; // the column kernel reads the result next
        return launch_rows(dtype, l.row_n, kRowResidual, a, l.grid, s) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_HIP;
    }
    if (launch_residual(dtype, out, y, t, alpha, rp.total, s) != hipSuccess) return AFFT_ERROR_HIP;
    return run_plan(rp, -1, out, out, 1.0, s);
}

#define AFFT_SPEC_E(X) X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16)

template <typename Real> const void *colconv_fn(bool vjp, int E, int64_t h)
{
#define AFFT_COLCONV_CT(H, EV, ...)                                                               \
    if (h == H && E == EV)                                                                       \
        return vjp ? reinterpret_cast<const void *>(afft_colconv_vjp<Real, EV, H, __VA_ARGS__>)  \
                   : reinterpret_cast<const void *>(afft_colconv<Real, EV, H, __VA_ARGS__>);
    // compile-time transforms for the common heights (the radices of the row kernels)
    AFFT_COLCONV_CT(256, 16, 16, 16)
    AFFT_COLCONV_CT(512, 16, 16, 16, 2)
#undef AFFT_COLCONV_CT
    switch (E) {
#define X(EV)                                                                                    \
    case EV:                                                                                     \
        return vjp ? reinterpret_cast<const void *>(afft_colconv_vjp<Real, EV, 0>)              \
                   : reinterpret_cast<const void *>(afft_colconv<Real, EV, 0>);
        AFFT_SPEC_E(X)
#undef X
    default: return nullptr;
    }
}

template <typename Real>
hipError_t launch_colconv(bool vjp, int E, const ColConvArgs &a, int block, int64_t grid, size_t lds,
                          hipStream_t s)
{
    const void *fn = colconv_fn<Real>(vjp, E, a.h);
    if (!fn) return hipErrorInvalidValue;
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    void *args[] = {const_cast<ColConvArgs *>(&a)};
    return hipLaunchKernel(fn, dim3((unsigned)grid), dim3(block), args, lds, s);
}

struct ColShape {
    LdsLayout L;
    int C, block;
    size_t lds;
};

bool col_shape(int dtype, int64_t h, ColShape &cs, bool vjp)
{
    if (h > 16384 || !is_5smooth(h)) return false;
    if (!choose_layout(h, dtype, cs.L, 16, vjp ? vjp_block(dtype == AFFT_C128) : -1)) return false;
    const int T = (int)(h / cs.L.E);
    const int maxb = vjp ? vjp_block(dtype == AFFT_C128) : max_block_threads(dtype == AFFT_C128, cs.L.E);
    const size_t esz = dtype == AFFT_C64 ? 8 : 16;
    cs.C = 16;
    while (cs.C > 1 && (cs.C * T > maxb || (size_t)cs.C * lds_padded_len((int)h) * esz > 80 * 1024)) cs.C /= 2;
    if (cs.C * T > maxb) return false;
    cs.block = cs.C * T;
    cs.lds = cs.L.npass > 1 ? (size_t)cs.C * lds_padded_len((int)h) * esz : 0;
    return cs.lds <= 160 * 1024;
}

// per-(device, dtype, h) column tables, kept for the life of the process
struct ColTables {
    Plan owner;
    PassTable pt;
    const void *tw = nullptr;
    const void *pw = nullptr; // W_h^m, m in [0, h)
};

int col_tables(int dtype, int64_t h, int device, const ColShape &cs, const ColTables *&out)
{
    static std::mutex mu;
    static std::map<std::tuple<int, int, int64_t, int>, ColTables *> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(device, dtype, h, cs.L.E);
    auto it = cache.find(key);
    if (it != cache.end()) {
        out = it->second;
        return AFFT_SUCCESS;
    }
    DeviceInfo dev;
    int st = device_info(device, dev);
    if (st) return st;
    ColTables *ct = new ColTables();
    ct->owner.dtype = dtype;
    ct->owner.device = device;
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
        delete ct;
        return AFFT_ERROR_HIP;
    }
    st = table_pass_twiddles(ct->owner, dev, s, h, cs.L, ct->pt, ct->tw);
    if (!st) {
        void *pw = plan_alloc(ct->owner, dev, h * (int64_t)(dtype == AFFT_C64 ? 8 : 16), st);
        if (!st && gen_power_table(dtype, pw, h, 1, h, s) != hipSuccess) st = AFFT_ERROR_HIP;
        ct->pw = pw;
    }
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    if (st) {
        free_plan(ct->owner);
        delete ct;
        return st;
    }
    cache[key] = ct;
    out = ct;
    return AFFT_SUCCESS;
}

// Unfused composition for shapes without a fused column layout: batched 2D plans plus pointwise
// kernels.  Same mathematics, more array passes.
int conv_fwd_unfused(int dtype, int64_t B, int64_t h, int64_t w, const void *x, const void *H, void *y,
                     void *spec, int dev, hipStream_t s)
{
    const int64_t ext[3] = {B, h, w};
    const Plan *p2 = nullptr;
    int st = cached_plan(p2, dtype, 3, ext, AFFT_AXES_LAST2, dev);
    if (st) return st;
    if ((st = exec_plan(*p2, -1, x, spec ? spec : y, 1.0, s))) return st;
    if (pointwise_mul(dtype, y, spec ? spec : y, H, B * h * w, h * w, 0, 1.0, s) != hipSuccess) return AFFT_ERROR_HIP;
    return exec_plan(*p2, +1, y, y, 1.0 / (double)(h * w), s);
}

int conv_vjp_unfused(int dtype, int64_t B, int64_t h, int64_t w, const void *x, bool x_is_spec,
                     const void *H, const void *gy, void *gx, void *gH, int dev, hipStream_t s)
{
    const int64_t ext[3] = {B, h, w};
    const Plan *p2 = nullptr;
    int st = cached_plan(p2, dtype, 3, ext, AFFT_AXES_LAST2, dev);
    if (st) return st;
    const size_t esz = dtype == AFFT_C64 ? 8 : 16;
    const int64_t n = B * h * w, m = h * w;
    void *ws = nullptr;
    if (hipMallocAsync(&ws, 2 * (size_t)n * esz, s) != hipSuccess) return AFFT_ERROR_ALLOC;
    void *X = ws, *G = (char *)ws + (size_t)n * esz;
    const double scale = 1.0 / (double)m;
    if (!x_is_spec) st = exec_plan(*p2, -1, x, X, 1.0, s);
    if (!st) st = exec_plan(*p2, -1, gy, G, 1.0, s);
    if (!st && gH) {
        // X <- G * conj(F x), then the batch sum
        if (pointwise_mul(dtype, X, G, x_is_spec ? x : X, n, n, 1, 1.0, s) != hipSuccess) st = AFFT_ERROR_HIP;
        const int64_t g = std::min<int64_t>(4096, (m + 255) / 256);
        if (!st) {
            if (dtype == AFFT_C64)
                hipLaunchKernelGGL(afft_reduce_groups<float>, dim3((unsigned)g), dim3(256), 0, s, (float2 *)gH,
                                   (const float2 *)X, (int)B, m, (float)scale);
            else
                hipLaunchKernelGGL(afft_reduce_groups<double>, dim3((unsigned)g), dim3(256), 0, s, (double2 *)gH,
                                   (const double2 *)X, (int)B, m, scale);
            if (hipGetLastError() != hipSuccess) st = AFFT_ERROR_HIP;
        }
    }
    if (!st && gx) {
        if (pointwise_mul(dtype, G, G, H, n, m, 1, 1.0, s) != hipSuccess) st = AFFT_ERROR_HIP;
        if (!st) st = exec_plan(*p2, +1, G, gx, scale, s);
    }
    (void)hipFreeAsync(ws, s);
    return st;
}

int check_args(int dtype, int64_t B, int64_t h, int64_t w)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (B < 0 || h < 0 || w < 0) return AFFT_ERROR_INVALID_VALUE;
    return AFFT_SUCCESS;
}

} // namespace
} // namespace afft

using namespace afft;

static int conv_fwd(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x, const void *d_H, void *d_y,
                    void *d_spec, void *stream)
{
    int st = check_args(dtype, batch, h, w);
    if (st) return st;
    if (batch * h * w == 0) return AFFT_SUCCESS;
    if (!d_x || !d_H || !d_y) return AFFT_ERROR_INVALID_VALUE;
    if (d_spec && (d_spec == d_y || d_spec == d_H)) return AFFT_ERROR_INVALID_VALUE;
    int dev;
    if ((st = current_device(dev))) return st;
    ColShape cs;
    if (!col_shape(dtype, h, cs, false)) return conv_fwd_unfused(dtype, batch, h, w, d_x, d_H, d_y, d_spec, dev, (hipStream_t)stream);
    const ColTables *ct = nullptr;
    if ((st = col_tables(dtype, h, dev, cs, ct))) return st;
    const int64_t rows[2] = {batch * h, w};
    const Plan *rp = nullptr;
    if ((st = cached_plan(rp, dtype, 2, rows, AFFT_AXES_INNERMOST, dev))) return st;
    hipStream_t s = (hipStream_t)stream;
    // 1. forward row transforms x -> y
    if ((st = run_plan(*rp, -1, d_x, d_y, 1.0, stream))) return st;
    // 2. column transform, product with H, inverse column transform, in place on y
    ColConvArgs a{};
    a.data = d_y;
    a.spec = d_spec;
    a.H = d_H;
    a.tw = ct->tw;
    a.pt = ct->pt;
    a.pw = ct->pw;
    a.B = batch;
    a.h = h;
    a.w = w;
    a.C = cs.C;
    const int64_t grid = batch * ((w + cs.C - 1) / cs.C);
    hipError_t e = dtype == AFFT_C64 ? launch_colconv<float>(false, cs.L.E, a, cs.block, grid, cs.lds, s)
                                     : launch_colconv<double>(false, cs.L.E, a, cs.block, grid, cs.lds, s);
    if (e != hipSuccess) return AFFT_ERROR_HIP;
    // 3. inverse row transforms with the 1/(h w) scale
    return run_plan(*rp, +1, d_y, d_y, 1.0 / (double)(h * w), stream);
}

extern "C" int afft_spectral_conv_fwd(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                      const void *d_H, void *d_y, void *stream)
{
    return conv_fwd(dtype, batch, h, w, d_x, d_H, d_y, nullptr, stream);
}

extern "C" int afft_spectral_conv_fwd_save(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                           const void *d_H, void *d_y, void *d_spec, void *stream)
{
    if (!d_spec && batch * h * w != 0) return AFFT_ERROR_INVALID_VALUE;
    return conv_fwd(dtype, batch, h, w, d_x, d_H, d_y, d_spec, stream);
}

// d_x is x, or F x when x_is_spec.  The upstream gradient is d_gy, or (residual form, d_t not
// null) alpha (d_y - d_t), formed as the first row pass loads y and t.
static int conv_vjp(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x, bool x_is_spec,
                    const void *d_H, const void *d_gy, void *d_gx, void *d_gH, void *stream,
                    const void *d_y = nullptr, const void *d_t = nullptr, double alpha = 0.0)
{
    int st = check_args(dtype, batch, h, w);
    if (st) return st;
    const size_t esz = dtype == AFFT_C64 ? 8 : 16;
    hipStream_t s = (hipStream_t)stream;
    const bool residual = d_t != nullptr;
    if (h * w == 0) return AFFT_SUCCESS;
    if (!d_x || !d_H || (residual ? !d_y : !d_gy) || (!d_gx && !d_gH)) return AFFT_ERROR_INVALID_VALUE;
    if (x_is_spec && (d_x == d_gx || d_x == d_gH)) return AFFT_ERROR_INVALID_VALUE;
    if (batch == 0) {
        if (d_gH && hipMemsetAsync(d_gH, 0, (size_t)(h * w) * esz, s) != hipSuccess) return AFFT_ERROR_HIP;
        return AFFT_SUCCESS;
    }
    int dev;
    if ((st = current_device(dev))) return st;
    ColShape cs;
    if (!col_shape(dtype, h, cs, true)) {
        if (!residual)
            return conv_vjp_unfused(dtype, batch, h, w, d_x, x_is_spec, d_H, d_gy, d_gx, d_gH, dev, s);
        // the unfused path takes the upstream gradient stored
        void *gy = nullptr;
        const int64_t n = batch * h * w;
        if (hipMallocAsync(&gy, (size_t)n * esz, s) != hipSuccess) return AFFT_ERROR_ALLOC;
        if (launch_residual(dtype, gy, d_y, d_t, alpha, n, s) != hipSuccess) st = AFFT_ERROR_HIP;
        if (!st) st = conv_vjp_unfused(dtype, batch, h, w, d_x, x_is_spec, d_H, gy, d_gx, d_gH, dev, s);
        (void)hipFreeAsync(gy, s);
        return st;
    }
    const ColTables *ct = nullptr;
    if ((st = col_tables(dtype, h, dev, cs, ct))) return st;
    const int64_t rows[2] = {batch * h, w};
    const Plan *rp = nullptr;
    if ((st = cached_plan(rp, dtype, 2, rows, AFFT_AXES_INNERMOST, dev))) return st;
    DeviceInfo di;
    if ((st = device_info(dev, di))) return st;
    // groups of images per column tile: enough workgroups to cover the chip once (one workgroup of
    // this kernel fits a CU).  Each group's gH partial is one more h x w array written and read
    // back: on 64 x 512^2, 8 groups took the squared-error VJP 303-308 us, 16 groups 303-315 us,
    // 32 groups 315-319 us and 64 groups 346-354 us
    const int64_t tiles = (w + cs.C - 1) / cs.C;
    int groups = (int)std::min<int64_t>(batch, std::max<int64_t>(1, (di.cus + tiles - 1) / tiles));
    const size_t arr = (size_t)(batch * h * w) * esz;
    const size_t part = (size_t)groups * (size_t)(h * w) * esz;
    void *ws = nullptr;
    // row-transformed x (unless F x is given); row-transformed gy lives in gx (or in the workspace
    // when gx is not wanted)
    const size_t xr_bytes = x_is_spec ? 0 : arr;
    const size_t ws_bytes = xr_bytes + part + (d_gx ? 0 : arr);
    if (hipMallocAsync(&ws, ws_bytes, s) != hipSuccess) return AFFT_ERROR_ALLOC;
    char *base = (char *)ws;
    const void *xr = x_is_spec ? d_x : (const void *)base;
    void *partial = base + xr_bytes;
    void *gr = d_gx ? d_gx : (void *)(base + xr_bytes + part);
    if (!x_is_spec) st = run_plan(*rp, -1, d_x, const_cast<void *>(xr), 1.0, stream);
    if (!st) st = residual ? residual_rows(*rp, dtype, d_y, d_t, alpha, gr, s) : run_plan(*rp, -1, d_gy, gr, 1.0, stream);
    if (!st) {
        ColConvArgs a{};
        a.data = gr;
        a.other = xr;
        a.other_is_spec = x_is_spec ? 1 : 0;
        a.H = d_H;
        a.partial = partial;
        a.tw = ct->tw;
        a.pt = ct->pt;
        a.pw = ct->pw;
        a.B = batch;
        a.h = h;
        a.w = w;
        a.C = cs.C;
        a.groups = groups;
            const int64_t grid = (int64_t)groups * tiles;
        hipError_t e = dtype == AFFT_C64 ? launch_colconv<float>(true, cs.L.E, a, cs.block, grid, cs.lds, s)
                                         : launch_colconv<double>(true, cs.L.E, a, cs.block, grid, cs.lds, s);
        if (e != hipSuccess) st = AFFT_ERROR_HIP;
    }
    const double scale = 1.0 / (double)(h * w);
    if (!st && d_gH) {
        const int64_t n = h * w;
        const int64_t g = std::min<int64_t>(4096, (n + 255) / 256);
        if (dtype == AFFT_C64)
            hipLaunchKernelGGL(afft_reduce_groups<float>, dim3((unsigned)g), dim3(256), 0, s, (float2 *)d_gH,
                               (const float2 *)partial, groups, n, (float)scale);
        else
            hipLaunchKernelGGL(afft_reduce_groups<double>, dim3((unsigned)g), dim3(256), 0, s, (double2 *)d_gH,
                               (const double2 *)partial, groups, n, scale);
        if (hipGetLastError() != hipSuccess) st = AFFT_ERROR_HIP;
    }
    if (!st && d_gx) st = run_plan(*rp, +1, d_gx, d_gx, scale, stream);
    (void)hipFreeAsync(ws, s);
    return st;
}

extern "C" int afft_spectral_conv_vjp(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                      const void *d_H, const void *d_gy, void *d_gx, void *d_gH,
                                      void *stream)
{
    return conv_vjp(dtype, batch, h, w, d_x, false, d_H, d_gy, d_gx, d_gH, stream);
}

extern "C" int afft_spectral_conv_vjp_residual(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_x,
                                               int x_is_spectrum, const void *d_H, const void *d_y,
                                               const void *d_t, double alpha, void *d_gx, void *d_gH, void *stream)
{
    if (!d_t && batch * h * w != 0) return AFFT_ERROR_INVALID_VALUE;
    return conv_vjp(dtype, batch, h, w, d_x, x_is_spectrum != 0, d_H, nullptr, d_gx, d_gH, stream, d_y, d_t, alpha);
}

extern "C" int afft_spectral_conv_vjp_saved(int dtype, int64_t batch, int64_t h, int64_t w, const void *d_spec,
                                            const void *d_H, const void *d_gy, void *d_gx, void *d_gH,
                                            void *stream)
{
    return conv_vjp(dtype, batch, h, w, d_spec, true, d_H, d_gy, d_gx, d_gH, stream);
}
