// afft_cols.hip -- register-resident transforms for the passes of a strided axis, and the
// batched transpose that ends a six-step transform.
//
//   afft_regcol<Real, R>   one thread owns one whole length-R transform of a strided batch:
//                          R loads, an in-register DFT_R (dft_split for R = 32, 64), the optional
//                          four-step twiddle, R stores.  No LDS and no barrier.
//
// Transform j is the thread's global index and the innermost level of its address is the
// column, so the 64 lanes of a wave own 64 neighbouring columns: every load and store
// instruction of a wave touches one contiguous run of 64 elements (512 B in FP32, 1 KiB in
// FP64).  A strided axis too long for a line-wide single-workgroup tile (4096 columns of 4096
// in FP32 would need 512 KiB of LDS for 16 columns) is split four-step into passes of these
// lengths, each one HBM read and one HBM write of the array with whole lines.
#include <hip/hip_runtime.h>

#include <utility>

#include "../../include/afft.h"
#include "afft_device.hpp"
#include "afft_internal.hpp"

namespace afft {

constexpr int kRegColBlock = 256;

template <typename Real, int R>
__global__ void __launch_bounds__(kRegColBlock) __attribute__((amdgpu_waves_per_eu(2, 8)))
afft_regcol(const FftArgs a)
{
    using V = typename cplx_of<Real>::type;
    const int64_t tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t j = tile * kRegColBlock + threadIdx.x;
    if (j >= a.ntrans) return;

    const int64_t ibase = addr_base(a.ain, j);
    const V *in = reinterpret_cast<const V *>(a.in) + ibase;
    V reg[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const V v = c_load(in + (int64_t)i * a.ain.es, (bool)a.nt_in);
        reg[i] = a.conj_in ? c_conj(v) : v;
    }
    if (a.sg_in.on) { // apart from the loads, which stay one batch of independent instructions
#pragma unroll
        for (int i = 0; i < R; ++i) reg[i] = sign_apply(a.sg_in, ibase + (int64_t)i * a.ain.es, reg[i]);
    }
    dft<R>::run(reg);

    if (a.post_L) {
        // four-step twiddle W_L^{jl k}, k = k0 + R0 k1: W^{jl k0} (R0 lookups) times
        // W^{R0 jl k1} (R / R0 lookups); jl < L / R and k < R keep every exponent below L
        constexpr int R0 = R >= 8 ? 8 : R;
        const V *hi = reinterpret_cast<const V *>(a.post_hi);
        const V *lo = reinterpret_cast<const V *>(a.post_lo);
        const int64_t jl = addr_level(a.ain, j, a.post_level);
        V w0[R0];
#pragma unroll
        for (int k0 = 0; k0 < R0; ++k0) w0[k0] = power_lookup(hi, lo, jl * k0, a.post_shift);
#pragma unroll
        for (int k1 = 0; k1 < R / R0; ++k1) {
            const V w1 = power_lookup(hi, lo, jl * (int64_t)(R0 * k1), a.post_shift);
#pragma unroll
            for (int k0 = 0; k0 < R0; ++k0) reg[k0 + R0 * k1] = c_mul(reg[k0 + R0 * k1], c_mul(w0[k0], w1));
        }
    }

    const int64_t obase = addr_base(a.aout, j);
    V *out = reinterpret_cast<V *>(a.out) + obase;
    const Real sc = (Real)a.scale;
    if (a.sg_out.on) {
#pragma unroll
        for (int k = 0; k < R; ++k) reg[k] = sign_apply(a.sg_out, obase + (int64_t)k * a.aout.es, reg[k]);
    }
#pragma unroll
    for (int k = 0; k < R; ++k) {
        V v = reg[k];
        if (a.conj_out) v = c_conj(v);
        c_store(out + (int64_t)k * a.aout.es, c_scale(v, sc), (bool)a.nt_out);
    }
}

// ----------------------------------------------------------------------------------------------
// The same transforms with each column spread over Q = 2 or 4 neighbouring lanes of a quad
// (afft_regcol_q): Q times the waves of afft_regcol, each holding R/Q elements, so loads of later
// waves overlap the arithmetic and stores of earlier ones.  With M = R/Q, lane q of a column holds
// x[q + Q m] (m < M), and the transform is the two-level split of index i = q + Q m,
// k = a + M b:
//   Y_q[a] = W_R^{q a} * DFT_M(x[q + Q m])[a]                     (in the lane's registers)
//   X[a + M b] = sum_q W_Q^{q b} Y_q[a]                            (across the Q lanes)
// The cross-lane step exchanges Q x Q blocks of Y (a = Q s + r) with DPP quad permutations, so
// that lane l holds Y_q[Q s + l] for every q, and then runs DFT_Q in registers: lane l ends
// with X[Q s + l + M b] for s < M/Q, b < Q.  The lanes of a wave still read and write whole
// lines: a wave instruction covers 64/Q columns on Q neighbouring rows.

// lane l receives the value lane l ^ X of its quad holds (DPP quad_perm, a VALU operation)
template <int X> __device__ __forceinline__ int quad_xor_i(int v)
{
    static_assert(X == 1 || X == 2, "quad lanes");
    // quad_perm selectors: lane i of each quad reads lane ((ctrl >> 2i) & 3)
    constexpr int ctrl = X == 1 ? 0xB1 /* 1,0,3,2 */ : 0x4E /* 2,3,0,1 */;
    return __builtin_amdgcn_update_dpp(0, v, ctrl, 0xF, 0xF, false);
}

template <int X> __device__ __forceinline__ float quad_xor(float v)
{
    return __builtin_bit_cast(float, quad_xor_i<X>(__builtin_bit_cast(int, v)));
}

template <int X> __device__ __forceinline__ double quad_xor(double v)
{
    const long long b = __builtin_bit_cast(long long, v);
    const int lo = quad_xor_i<X>((int)(b & 0xffffffffll));
    const int hi = quad_xor_i<X>((int)(b >> 32));
    return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int X, typename V> __device__ __forceinline__ V quad_xor_c(V v)
{
    return c_make<V>(quad_xor<X>(v.x), quad_xor<X>(v.y));
}

// Transpose of the Q x Q block (lane l of the quad group, registers y[0..Q-1]) across the lanes:
// after it lane l holds what lane r held in register l, in register r.  One stage per lane bit:
// element (l, r) with bit B of l and r different trades places with (l ^ B, r ^ B).
template <int Q, int B, typename V> __device__ __forceinline__ void quad_transpose_stage(V *y, int l)
{
    V p[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r) p[r] = quad_xor_c<B>(y[r ^ B]);
#pragma unroll
    for (int r = 0; r < Q; ++r)
        if (((l & B) != 0) != ((r & B) != 0)) y[r] = p[r];
}

template <int Q, typename V> __device__ __forceinline__ void quad_transpose(V *y, int l)
{
    quad_transpose_stage<Q, 1>(y, l);
    if constexpr (Q == 4) quad_transpose_stage<Q, 2>(y, l);
}

// v * W_R^{q a} for the lane's q < Q (a compile-time, q at run time)
template <int R, int Q, int A, typename V> __device__ __forceinline__ V lane_twiddle(V v, int q)
{
    const V t1 = twiddle_const<R, A>(v);
    if constexpr (Q == 2) return q ? t1 : v;
    const V t2 = twiddle_const<R, 2 * A>(v);
    const V t3 = twiddle_const<R, 3 * A>(v);
    return q == 0 ? v : q == 1 ? t1 : q == 2 ? t2 : t3;
}

template <int R, int Q, typename V, int... As>
__device__ __forceinline__ void lane_twiddles(V *y, int q, std::integer_sequence<int, As...>)
{
    ((y[As] = lane_twiddle<R, Q, As>(y[As], q)), ...);
}

template <typename Real, int R, int Q>
__global__ void __launch_bounds__(kRegColBlock) afft_regcol_q(const FftArgs a)
{
    using V = typename cplx_of<Real>::type;
    constexpr int M = R / Q;
    static_assert(M % Q == 0, "Q x Q blocks");
    const int64_t tile = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int q = threadIdx.x % Q;
    const int64_t j = tile * (kRegColBlock / Q) + threadIdx.x / Q;
    // the Q lanes of a column stay or leave together, so the quad exchanges below only ever read
    // lanes of the same column
    if (j >= a.ntrans) return;

    const int64_t ibase = addr_base(a.ain, j);
    const V *in = reinterpret_cast<const V *>(a.in) + ibase;
    V y[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const V v = c_load(in + (int64_t)(q + Q * m) * a.ain.es, (bool)a.nt_in);
        y[m] = a.conj_in ? c_conj(v) : v;
    }
    if (a.sg_in.on) {
#pragma unroll
        for (int m = 0; m < M; ++m) y[m] = sign_apply(a.sg_in, ibase + (int64_t)(q + Q * m) * a.ain.es, y[m]);
    }
    dft<M>::run(y);
    lane_twiddles<R, Q>(y, q, std::make_integer_sequence<int, M>{});
#pragma unroll
    for (int s = 0; s < M / Q; ++s) {
        quad_transpose<Q>(y + Q * s, q);
        dft<Q>::run(y + Q * s); // lane q: y[Q s + b] = X[Q s + q + M b]
    }

    // output k of register Q s + b
    auto kidx = [&](int s, int b) { return (int64_t)(Q * s + q + M * b); };
    if (a.post_L) {
        // four-step twiddle W_L^{jl k}, k = q + Q s + M b: W^{jl q} W^{jl Q s} W^{jl M b}
        const V *hi = reinterpret_cast<const V *>(a.post_hi);
        const V *lo = reinterpret_cast<const V *>(a.post_lo);
        const int64_t jl = addr_level(a.ain, j, a.post_level);
        const V wq = power_lookup(hi, lo, jl * q, a.post_shift);
        V wb[Q];
#pragma unroll
        for (int b = 0; b < Q; ++b) wb[b] = power_lookup(hi, lo, jl * (int64_t)(M * b), a.post_shift);
#pragma unroll
        for (int s = 0; s < M / Q; ++s) {
            const V ws = c_mul(wq, power_lookup(hi, lo, jl * (int64_t)(Q * s), a.post_shift));
#pragma unroll
            for (int b = 0; b < Q; ++b) y[Q * s + b] = c_mul(y[Q * s + b], c_mul(ws, wb[b]));
        }
    }

    const int64_t obase = addr_base(a.aout, j);
    V *out = reinterpret_cast<V *>(a.out) + obase;
    const Real sc = (Real)a.scale;
#pragma unroll
    for (int s = 0; s < M / Q; ++s) {
#pragma unroll
        for (int b = 0; b < Q; ++b) {
            V v = y[Q * s + b];
            if (a.sg_out.on) v = sign_apply(a.sg_out, obase + kidx(s, b) * a.aout.es, v);
            if (a.conj_out) v = c_conj(v);
            c_store(out + kidx(s, b) * a.aout.es, c_scale(v, sc), (bool)a.nt_out);
        }
    }
}

bool regcol_supported(int dtype, int64_t n)
{
    // FP64 stops at 32: 64 FP64 complex values alone would take 256 VGPRs
    return n == 16 || n == 32 || (n == 64 && dtype == AFFT_C64);
}

int regcol_block() { return kRegColBlock; }

template <typename Real>
static hipError_t launch_regcol_t(const FftArgs &a, int64_t grid, hipStream_t s)
{
    if (a.lanes == 2 || a.lanes == 4) {
        const int key = a.E * 8 + a.lanes;
        switch (key) {
        case 16 * 8 + 2: hipLaunchKernelGGL((afft_regcol_q<Real, 16, 2>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        case 16 * 8 + 4: hipLaunchKernelGGL((afft_regcol_q<Real, 16, 4>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        case 32 * 8 + 2: hipLaunchKernelGGL((afft_regcol_q<Real, 32, 2>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        case 32 * 8 + 4: hipLaunchKernelGGL((afft_regcol_q<Real, 32, 4>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        case 64 * 8 + 2: hipLaunchKernelGGL((afft_regcol_q<Real, 64, 2>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        case 64 * 8 + 4: hipLaunchKernelGGL((afft_regcol_q<Real, 64, 4>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
        default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    switch (a.E) {
    case 16: hipLaunchKernelGGL((afft_regcol<Real, 16>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
    case 32: hipLaunchKernelGGL((afft_regcol<Real, 32>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a); break;
    case 64:
        if constexpr (sizeof(Real) == 4) {
            hipLaunchKernelGGL((afft_regcol<Real, 64>), dim3((unsigned)grid), dim3(kRegColBlock), 0, s, a);
            break;
        }
        return hipErrorInvalidValue;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_regcol(int dtype, const FftArgs &a, int64_t grid, hipStream_t s)
{
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    return dtype == AFFT_C64 ? launch_regcol_t<float>(a, grid, s) : launch_regcol_t<double>(a, grid, s);
}


// ----------------------------------------------------------------------------------------------
// Batched strided transpose, the last pass of a six-step transform: matrix o = o1 + nb1 * o2
// (rows x cols, element (r, c) at in + o1 in_bs1 + o2 in_bs2 + r in_rs + c) is written transposed,
// element (c, r) at out + o1 out_bs1 + (toff ? toff[o2] : o2 out_bs2) + c out_rs + r.  A square
// tile goes through LDS so that both the loads (along a row of the input) and the stores (along
// a row of the output) are contiguous runs of TILE elements; the LDS row pitch TILE + 1 keeps
// the column-wise LDS reads off a single bank.

template <typename Real, int TILE>
__global__ void __launch_bounds__(256) afft_transpose(const TransposeArgs a)
{
    using V = typename cplx_of<Real>::type;
    constexpr int RS = 256 / TILE; // tile rows per sweep
    __shared__ V tile[TILE][TILE + 1];
    const int64_t b = a.xcd_remap ? xcd_tile(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
    const int64_t ntx = (a.cols + TILE - 1) / TILE, nty = (a.rows + TILE - 1) / TILE;
    const int64_t bx = b % ntx, q = b / ntx;
    const int64_t by = q % nty, o = q / nty;
    if (o >= a.nbatch) return; // whole workgroup: no barrier is skipped by part of it
    const int64_t o1 = o % a.nb1, o2 = o / a.nb1;
    const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
    const V *in = reinterpret_cast<const V *>(a.in) + o1 * a.in_bs1 + o2 * a.in_bs2;
    const int64_t obase = o1 * a.out_bs1 + (a.toff ? a.toff[o2] : o2 * a.out_bs2);
    V *out = reinterpret_cast<V *>(a.out) + obase;
#pragma unroll
    for (int i = 0; i < TILE / RS; ++i) {
        const int64_t r = by * TILE + ty + i * RS, c = bx * TILE + tx;
        if (r < a.rows && c < a.cols) tile[ty + i * RS][tx] = in[r * a.in_rs + c];
    }
    __syncthreads();
    const Real sc = (Real)a.scale;
    const bool sign = a.sg_out.on;
#pragma unroll
    for (int i = 0; i < TILE / RS; ++i) {
        const int64_t c = bx * TILE + ty + i * RS, r = by * TILE + tx;
        if (r < a.rows && c < a.cols) {
            V v = tile[tx][ty + i * RS];
            if (a.conj_out) v = c_conj(v);
            v = c_scale(v, sc);
            if (sign) v = sign_apply(a.sg_out, obase + c * a.out_rs + r, v);
            c_store(out + c * a.out_rs + r, v, (bool)a.nt_out);
        }
    }
}

int64_t transpose_grid(int dtype, int64_t rows, int64_t cols, int64_t nbatch)
{
    const int64_t T = dtype == AFFT_C64 ? 64 : 32;
    return ((cols + T - 1) / T) * ((rows + T - 1) / T) * nbatch;
}

hipError_t launch_transpose(int dtype, const TransposeArgs &a, hipStream_t s)
{
    const int64_t grid = transpose_grid(dtype, a.rows, a.cols, a.nbatch);
    if (grid <= 0) return hipSuccess;
    if (grid > 0x7fffffff) return hipErrorInvalidValue;
    if (dtype == AFFT_C64)
        hipLaunchKernelGGL((afft_transpose<float, 64>), dim3((unsigned)grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((afft_transpose<double, 32>), dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace afft
