// afft_cols.hip -- register-resident transforms for the passes of a strided axis.
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

    const V *in = reinterpret_cast<const V *>(a.in) + addr_base(a.ain, j);
    V reg[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        const V v = c_load(in + (int64_t)i * a.ain.es, (bool)a.nt_in);
        reg[i] = a.conj_in ? c_conj(v) : v;
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

    V *out = reinterpret_cast<V *>(a.out) + addr_base(a.aout, j);
    const Real sc = (Real)a.scale;
#pragma unroll
    for (int k = 0; k < R; ++k) {
        V v = reg[k];
        if (a.conj_out) v = c_conj(v);
        c_store(out + (int64_t)k * a.aout.es, c_scale(v, sc), (bool)a.nt_out);
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

} // namespace afft
