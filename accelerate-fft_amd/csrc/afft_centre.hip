// afft_centre.hip -- the centring and shifting transforms of DFT/Centre.hs, standalone and fused
// into a transform.
//
//   centreND x  = x * (-1)^(sum of the multi-index)                       (Centre.hs:36-65)
//   shiftND x   = backpermute, index i_a -> (i_a + ceil(n_a/2)) mod n_a   (Centre.hs:70-82, 104-145)
//   ishiftND x  = backpermute, index i_a -> (i_a + floor(n_a/2)) mod n_a  (Centre.hs:84-101, 125-164)
//
// afft_fft_centred computes post(op(mode, pre(x))) for pre, post in {none, centre, shift, ishift}
// with the sign of a centring folded into the transform's first load (pre) or last store (post).
// A roll next to a transform is itself a sign on the other side when every extent is even:
//   shift(F x) = ishift(F x) = F(x * (-1)^(sum m))      F(shift x) = F(ishift x) = (-1)^(sum k) * F x
// (the roll by n/2 multiplies by e^(+-i pi k) = (-1)^k), so with even extents every combination
// runs as the transform alone, with no extra pass; a roll over an odd extent is a separate
// afft_reorder pass through a temporary.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/afft.h"
#include "afft_device.hpp"
#include "afft_internal.hpp"
#include "afft_plan.hpp"

namespace afft {

struct ReorderArgs {
    const void *in;
    void *out;
    int64_t total;
    SignSpec sg;        // coordinates of an output offset (and its sign when `centre`)
    int centre;
    int64_t roll[3];    // source coordinate = (coordinate + roll[a]) mod ext[a]
    int64_t stride[3];
};

template <typename Real>
__global__ void __launch_bounds__(256) afft_reorder(const ReorderArgs a)
{
    using V = typename cplx_of<Real>::type;
    const V *in = reinterpret_cast<const V *>(a.in);
    V *out = reinterpret_cast<V *>(a.out);
    for (int64_t off = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; off < a.total;
         off += (int64_t)gridDim.x * blockDim.x) {
        // coordinates from the fast dividers: coord_a = q_a - ext_a * q_{a-1}
        int64_t src = 0;
        int parity = 0;
        uint64_t qprev = 0;
        for (int d = 0; d < a.sg.rank; ++d) {
            const uint64_t q = ((uint64_t)off * a.sg.div[d].m) >> a.sg.div[d].sh;
            const int64_t c = (int64_t)(q - (uint64_t)a.sg.ext[d] * qprev);
            int64_t r = c + a.roll[d];
            if (r >= a.sg.ext[d]) r -= a.sg.ext[d];
            src += r * a.stride[d];
            parity ^= (int)(c & 1);
            qprev = q;
        }
        V v = in[src];
        if (a.centre && parity) v = c_make<V>(-v.x, -v.y);
        out[off] = v;
    }
}

static hipError_t launch_reorder(int dtype, const ReorderArgs &a, hipStream_t s)
{
    if (a.total <= 0) return hipSuccess;
    const int64_t grid = std::min<int64_t>((a.total + 255) / 256, 65536);
    if (dtype == AFFT_C64)
        hipLaunchKernelGGL(afft_reorder<float>, dim3((unsigned)grid), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL(afft_reorder<double>, dim3((unsigned)grid), dim3(256), 0, s, a);
    return hipGetLastError();
}

// the reorder pass for one of centre / shift / ishift, or false if `which` is none
static bool reorder_args(int which, int rank, const int64_t *ext, ReorderArgs &a)
{
    a = ReorderArgs{};
    if (which == AFFT_REORDER_NONE) return false;
    make_sign_spec(rank, ext, a.sg);
    a.total = 1;
    for (int d = 0; d < rank; ++d) a.total *= ext[d];
    int64_t stride = a.total;
    for (int d = 0; d < rank; ++d) {
        stride /= ext[d];
        a.stride[d] = stride;
        a.roll[d] = which == AFFT_SHIFT ? (ext[d] / 2 + (ext[d] & 1)) : which == AFFT_ISHIFT ? ext[d] / 2 : 0;
    }
    a.centre = which == AFFT_CENTRE;
    return true;
}

static int check_reorder_shape(int dtype, int rank, const int64_t *extents)
{
    if (dtype != AFFT_C64 && dtype != AFFT_C128) return AFFT_ERROR_UNSUPPORTED_DTYPE;
    if (rank < 1 || rank > 3) return AFFT_ERROR_UNSUPPORTED_RANK;
    if (!extents) return AFFT_ERROR_INVALID_VALUE;
    int64_t total = 1;
    for (int d = 0; d < rank; ++d) {
        if (extents[d] < 0) return AFFT_ERROR_INVALID_VALUE;
        total *= extents[d];
    }
    // the index arithmetic runs on 31-bit offsets
    if (total >= ((int64_t)1 << 31)) return AFFT_ERROR_UNSUPPORTED_SIZE;
    return AFFT_SUCCESS;
}

static bool valid_which(int w) { return w >= AFFT_REORDER_NONE && w <= AFFT_ISHIFT; }

} // namespace afft

using namespace afft;

extern "C" int afft_reorder(int which, int dtype, int rank, const int64_t *extents, const void *d_in,
                            void *d_out, void *stream)
{
    if (!valid_which(which)) return AFFT_ERROR_INVALID_VALUE;
    int st = check_reorder_shape(dtype, rank, extents);
    if (st) return st;
    ReorderArgs a;
    int64_t total = 1;
    for (int d = 0; d < rank; ++d) total *= extents[d];
    if (total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    // an element's source is another element: in place would read overwritten values
    if (d_in == d_out && which != AFFT_CENTRE && which != AFFT_REORDER_NONE) return AFFT_ERROR_INVALID_VALUE;
    const size_t bytes = (size_t)total * (dtype == AFFT_C64 ? 8 : 16);
    if (!reorder_args(which, rank, extents, a))
        return d_in == d_out || hipMemcpyAsync(d_out, d_in, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess
                   ? AFFT_SUCCESS : AFFT_ERROR_HIP;
    a.in = d_in;
    a.out = d_out;
    return launch_reorder(dtype, a, (hipStream_t)stream) == hipSuccess ? AFFT_SUCCESS : AFFT_ERROR_HIP;
}

extern "C" int afft_fft_centred(int mode, int dtype, int rank, const int64_t *extents, int pre, int post,
                                const void *d_in, void *d_out, void *stream)
{
    if (mode != AFFT_FORWARD && mode != AFFT_REVERSE && mode != AFFT_INVERSE) return AFFT_ERROR_INVALID_VALUE;
    if (!valid_which(pre) || !valid_which(post)) return AFFT_ERROR_INVALID_VALUE;
    int st = check_reorder_shape(dtype, rank, extents);
    if (st) return st;
    int64_t total = 1;
    bool all_even = true;
    for (int d = 0; d < rank; ++d) {
        total *= extents[d];
        all_even = all_even && (extents[d] % 2 == 0);
    }
    if (total == 0) return AFFT_SUCCESS;
    if (!d_in || !d_out) return AFFT_ERROR_INVALID_VALUE;
    int dev;
    if ((st = current_device(dev))) return st;
    const Plan *p = nullptr; // planned for the current device
    if ((st = cached_plan(p, dtype, rank, extents, AFFT_AXES_ALL, dev))) return st;
    hipStream_t s = (hipStream_t)stream;
    // fold every centring, and every roll over even extents, into a sign on the load or store
    bool load_sign = pre == AFFT_CENTRE, store_sign = post == AFFT_CENTRE;
    bool roll_pre = pre == AFFT_SHIFT || pre == AFFT_ISHIFT;
    bool roll_post = post == AFFT_SHIFT || post == AFFT_ISHIFT;
    if (roll_post && all_even) {
        load_sign = !load_sign;
        roll_post = false;
    }
    if (roll_pre && all_even) {
        store_sign = !store_sign;
        roll_pre = false;
    }
    SignSpec sg;
    make_sign_spec(rank, extents, sg);
    const size_t bytes = (size_t)total * (dtype == AFFT_C64 ? 8 : 16);
    void *tmp = nullptr;
    if ((roll_pre || roll_post) && hipMallocAsync(&tmp, bytes, s) != hipSuccess) return AFFT_ERROR_ALLOC;
    const void *src = d_in;
    void *dst = d_out;
    ReorderArgs ra;
    if (roll_pre) {
        reorder_args(pre, rank, extents, ra);
        ra.in = d_in;
        ra.out = tmp;
        if (launch_reorder(dtype, ra, s) != hipSuccess) st = AFFT_ERROR_HIP;
        src = tmp;
    }
    if (roll_post) dst = tmp; // with roll_pre too: the transform runs in place on the temporary
    const int dir = mode == AFFT_FORWARD ? -1 : 1;
    const double scale = mode == AFFT_INVERSE ? 1.0 / (double)total : 1.0;
    if (!st) st = exec_plan(*p, dir, src, dst, scale, s, nullptr, nullptr, load_sign ? &sg : nullptr,
                            store_sign ? &sg : nullptr);
    if (!st && roll_post) {
        reorder_args(post, rank, extents, ra);
        ra.in = tmp;
        ra.out = d_out;
        if (launch_reorder(dtype, ra, s) != hipSuccess) st = AFFT_ERROR_HIP;
    }
    if (tmp) (void)hipFreeAsync(tmp, s);
    return st;
}
