// afft_dft.hip -- Data.Array.Accelerate.Math.DFT on the device: the O(N^2) transforms dft, idft,
// dftG and dftGS (DFT.hs:42-111) with the roots of unity of DFT/Roots.hs:26-51.
//
// The reference's dftG extends the input to (sh :. shapeSize sh), reads `arr !! i` for EVERY
// linear index i of the array and folds that axis away (DFT.hs:68-92), so that
//     out[ix] = sum_i arr[i] * roots[sh_i :. (k * n_i) mod l],   (sh_i :. n_i) = fromIndex sh i,
// with k the innermost index of ix: for a vector the DFT, for a higher rank a value that depends
// only on k (every row receives the same sums over all rows).  The kernel reproduces exactly that:
// one workgroup per frequency k reduces over all elements, accumulating in FP64 with the index
// product (k * n_i) reduced in integers, and writes the value to every row.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/afft.h"
#include "afft_device.hpp"
#include "afft_internal.hpp"
#include "afft_plan.hpp"

namespace afft {

struct DftArgs {
    const void *in;
    const void *roots;  // full-shape roots (dftG), or null: the standard roots below
    int sign;           // standard roots: -1 rootsOfUnity (cos, -sin), +1 inverseRootsOfUnity
    void *out;
    int64_t total, l;   // elements, innermost extent
    int64_t k0, nk;     // frequencies k0 .. k0 + nk - 1
    int broadcast;      // 1: write every row's element k; 0: write out[k - k0] (dftGS)
    double scale;
};

constexpr int kDftBlock = 256;

template <typename Real>
__global__ void __launch_bounds__(kDftBlock) afft_dft_brute(const DftArgs a)
{
    using V = typename cplx_of<Real>::type;
    __shared__ double red[2][kDftBlock];
    const int64_t k = a.k0 + blockIdx.x;
    const V *x = reinterpret_cast<const V *>(a.in);
    const V *roots = reinterpret_cast<const V *>(a.roots);
    double sr = 0, si = 0;
    for (int64_t i = threadIdx.x; i < a.total; i += kDftBlock) {
        const int64_t row = i / a.l, n = i - row * a.l;
        const int64_t e = (k * n) % a.l; // k, n < l < 2^31: the product fits
        double wr, wi;
        if (roots) {
            const V w = roots[row * a.l + e];
            wr = (double)w.x;
            wi = (double)w.y;
        } else {
            // cos(2 pi e / l) + sign * i sin(2 pi e / l), from the reduced integer index
            sincospi(2.0 * (double)e / (double)a.l, &wi, &wr);
            wi *= a.sign;
        }
        const V v = x[i];
        sr += (double)v.x * wr - (double)v.y * wi;
        si += (double)v.x * wi + (double)v.y * wr;
    }
    red[0][threadIdx.x] = sr;
    red[1][threadIdx.x] = si;
    __syncthreads();
    for (int w = kDftBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            red[0][threadIdx.x] += red[0][threadIdx.x + w];
            red[1][threadIdx.x] += red[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    const V r = c_make<V>((Real)(red[0][0] * a.scale), (Real)(red[1][0] * a.scale));
    V *out = reinterpret_cast<V *>(a.out);
    if (a.broadcast) {
        for (int64_t row = threadIdx.x; row < a.total / a.l; row += kDftBlock) out[row * a.l + k] = r;
    } else if (threadIdx.x == 0) {
        out[k - a.k0] = r;
    }
}

static hipError_t launch_dft(int dtype, const DftArgs &a, hipStream_t s)
{
    if (a.nk <= 0) return hipSuccess;
    if (a.nk > 0x7fffffff) return hipErrorInvalidValue;
    if (dtype == AFFT_C64)
        hipLaunchKernelGGL(afft_dft_brute<float>, dim3((unsigned)a.nk), dim3(kDftBlock), 0, s, a);
    else
        hipLaunchKernelGGL(afft_dft_brute<double>, dim3((unsigned)a.nk), dim3(kDftBlock), 0, s, a);
    return hipGetLastError();
}

static int dft_shape(int dtype, int rank, const int64_t *extents, int64_t &total, int64_t &l)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (rank < 1 || rank > 8) return AFFT_ERROR_UNSUPPORTED_RANK;
    if (!extents) return AFFT_ERROR_INVALID_VALUE;
    total = 1;
    for (int d = 0; d < rank; ++d) {
        if (extents[d] < 0) return AFFT_ERROR_INVALID_VALUE;
        total *= extents[d];
    }
    l = extents[rank - 1];
    if (l >= ((int64_t)1 << 31)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    return AFFT_SUCCESS;
}

} // namespace afft

using namespace afft;

extern "C" int afft_dft(int mode, int dtype, int rank, const int64_t *extents, const void *d_in, void *d_out,
                        void *stream)
{
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    int64_t total, l;
    int st = dft_shape(dtype, rank, extents, total, l);
    if (st) return st;
    if (total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out || d_in == d_out) return AFFT_ERROR_INVALID_VALUE; // every output reads every input
    DftArgs a{};
    a.in = d_in;
    a.out = d_out;
    a.sign = mode == AFFT_FORWARD ? -1 : 1;
    a.total = total;
    a.l = l;
    a.k0 = 0;
    a.nk = l;
    a.broadcast = 1;
    a.scale = mode == AFFT_INVERSE ? 1.0 / (double)l : 1.0; // idft divides by n (DFT.hs:54-60)
    return launch_dft(dtype, a, (hipStream_t)stream) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_HIP;
}

extern "C" int afft_dftG(int dtype, int rank, const int64_t *extents, const void *d_roots, const void *d_in,
                         void *d_out, void *stream)
{
    int64_t total, l;
    int st = dft_shape(dtype, rank, extents, total, l);
    if (st) return st;
    if (total == 0) return AFFT_SUCCESS;
    if (!d_roots || !d_in || !d_out || d_in == d_out || d_roots == d_out) return AFFT_ERROR_INVALID_VALUE;
    DftArgs a{};
    a.in = d_in;
    a.roots = d_roots;
    a.out = d_out;
    a.total = total;
    a.l = l;
    a.k0 = 0;
    a.nk = l;
    a.broadcast = 1;
    a.scale = 1.0;
    return launch_dft(dtype, a, (hipStream_t)stream) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_HIP;
}

extern "C" int afft_dftGS(int dtype, int rank, const int64_t *extents, const int64_t *index, const void *d_roots,
                          const void *d_in, void *d_out, void *stream)
{
    int64_t total, l;
    int st = dft_shape(dtype, rank, extents, total, l);
    if (st) return st;
    if (!index || !d_roots || !d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    for (int d = 0; d < rank; ++d)
        if (index[d] < 0 || index[d] >= extents[d]) return AFFT_ERROR_INVALID_VALUE;
    if (total == 0) return AFFT_SUCCESS;
    DftArgs a{};
    a.in = d_in;
    a.roots = d_roots;
    a.out = d_out;
    a.total = total;
    a.l = l;
    a.k0 = index[rank - 1]; // only the innermost index enters the sum (DFT.hs:101-111)
    a.nk = 1;
    a.broadcast = 0;
    a.scale = 1.0;
    return launch_dft(dtype, a, (hipStream_t)stream) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_HIP;
}
