// afft_ext.cpp -- extension entry points with no counterpart in the reference API:
// the per-transform VJP, plus the entry points whose engines live in their own files.
#include <hip/hip_runtime.h>

#include "../../include/afft.h"
#include "afft_plan.hpp"

using namespace afft;

// y = op(mode, x) is linear: y = s * F_sign x with F_sign the unnormalised DFT of sign -1/+1
// over the op's axes.  Its adjoint (the VJP for a real loss, conjugate-Wirtinger convention) is
// s * F_sign^H = s * F_{-sign}: Forward's VJP is Reverse, Reverse's is Forward, and Inverse
// (Reverse / N, FFT.hs:83,110,141,172) has VJP Forward / N.
extern "C" int afft_fft_vjp(int op, int mode, int dtype, int rank, const int64_t *extents,
                            const void *d_gy, void *d_gx, void *stream)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    if (!extents) return AFFT_ERROR_INVALID_VALUE;
    int axes;
    switch (op) {
    case 0:
        if (rank < 1 || rank > 8) return AFFT_ERROR_UNSUPPORTED_RANK;
        axes = AFFT_AXES_INNERMOST;
        break;
    case 1:
    case 2:
    case 3:
        if (rank != op) return AFFT_ERROR_UNSUPPORTED_RANK;
        axes = AFFT_AXES_ALL;
        break;
    default: return AFFT_ERROR_INVALID_VALUE;
    }
    int64_t total = 1;
    for (int i = 0; i < rank; ++i) {
        if (extents[i] < 0) return AFFT_ERROR_INVALID_VALUE;
        total *= extents[i];
    }
    if (total == 0) return AFFT_SUCCESS;
    if (!d_gy || !d_gx) return AFFT_ERROR_INVALID_VALUE;
    const int64_t divisor = axes == AFFT_AXES_INNERMOST ? extents[rank - 1] : total;
    int dev, st;
    if ((st = current_device(dev))) return st;
    const Plan *p = nullptr;
    if ((st = cached_plan(p, dtype, rank, extents, axes, dev))) return st;
    const int dir = mode == AFFT_FORWARD ? +1 : -1;
    const double scale = mode == AFFT_INVERSE ? 1.0 / (double)divisor : 1.0;
    return run_plan(*p, dir, d_gy, d_gx, scale, stream);
}
