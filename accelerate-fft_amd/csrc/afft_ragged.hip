// afft_ragged.hip -- ragged batches: many 1D transforms of mixed lengths, stored back to back,
// in one persistent launch (kernel afft_ragged<Real, EMAX>).
//
// The host sorts the transforms into length classes and cuts each class into work items (a
// tile of same-length transforms, or one long transform), ordered by decreasing cost.  Each
// resident workgroup repeatedly takes the next item from a device-side ticket counter and runs
// it to completion; no workgroup ever waits for another, so the launch always drains.  The
// last workgroup to leave resets the counters, so the same launch can be replayed (as a captured
// hipGraph) without a memset node.
//
// Item kinds:
//   kDirect     C transforms of length L <= one workgroup's layout, Stockham passes in LDS
//   kBlue       C transforms of a non-5-smooth length with padded length M <= one workgroup:
//               the fused chirp-z transform (two forward passes sharing one LDS slot)
//   kTwoPhase   one long 5-smooth transform, L = L1 * L2, by the four-step scheme: L2 column
//               transforms of length L1 (twiddled) into the workgroup's scratch, then L1 column
//               transforms of length L2 into the output
//   kBlueTwo    one long non-5-smooth transform: chirp-z with both FFT_M done as kTwoPhase
//               steps through two scratch buffers
// Between phases the workgroup synchronises with an agent-scope fence and a barrier, so the
// scratch written by one wave is read back coherently by the others.
#include <hip/hip_runtime.h>

#include "afft_device.hpp"
#include "afft_ragged.hpp"

namespace afft {

__device__ __forceinline__ void phase_sync()
{
    __threadfence();
    __syncthreads();
}

// A batch of C same-length transforms at arbitrary element offsets.
template <typename Real, int E>
__device__ void run_direct(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const int L = K.L, T = L / E;
    const int t = threadIdx.x % T, c = threadIdx.x / T;
    const bool active = c < it.count;
    const int64_t off = active ? a.offsets[a.members[it.first + c]] : 0;
    const V *in = reinterpret_cast<const V *>(a.in);
    V *out = reinterpret_cast<V *>(a.out);
    V reg[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        V v = active ? in[off + t + i * T] : c_make<V>(0, 0);
        reg[i] = a.conj ? c_conj(v) : v;
    }
    V *slot = lds + (int64_t)(c < K.C ? c : 0) * lds_padded_len(L);
    fft_core<E>(reg, slot, 1, K.pt, reinterpret_cast<const V *>(K.tw), t, T, c < K.C);
    const Real sc = a.inverse ? (Real)(1.0 / (double)L) : (Real)1;
    if (active) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            V v = a.conj ? c_conj(reg[i]) : reg[i];
            out[off + t + i * T] = c_scale(v, sc);
        }
    }
}

template <typename Real, int E>
__device__ void run_blue(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const int n = K.L, M = K.M, T = M / E;
    const int t = threadIdx.x % T, c = threadIdx.x / T;
    const bool active = c < it.count;
    const int64_t off = active ? a.offsets[a.members[it.first + c]] : 0;
    const V *in = reinterpret_cast<const V *>(a.in);
    V *out = reinterpret_cast<V *>(a.out);
    const V *chirp = reinterpret_cast<const V *>(K.chirp);
    const V *filt = reinterpret_cast<const V *>(K.filt);
    const V *tw = reinterpret_cast<const V *>(K.tw);
    V reg[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = t + i * T;
        V v = c_make<V>(0, 0);
        if (active && m < n) {
            v = in[off + m];
            if (a.conj) v = c_conj(v);
            v = c_mul(v, chirp[m]);
        }
        reg[i] = v;
    }
    V *slot = lds + (int64_t)(c < K.C ? c : 0) * lds_padded_len(M);
    const bool on = c < K.C;
    fft_core<E>(reg, slot, 1, K.pt, tw, t, T, on);
#pragma unroll
    for (int i = 0; i < E; ++i) reg[i] = c_conj(c_mul(reg[i], on ? filt[t + i * T] : c_make<V>(0, 0)));
    fft_core<E>(reg, slot, 1, K.pt, tw, t, T, on);
    const Real sc = a.inverse ? (Real)(1.0 / (double)n) : (Real)1;
    if (active) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int m = t + i * T;
            if (m < n) {
                V v = c_mul(c_conj(reg[i]), chirp[m]);
                if (a.conj) v = c_conj(v);
                out[off + m] = c_scale(v, sc);
            }
        }
    }
}

// One column phase of the four-step scheme inside a workgroup: `ncol` transforms of length
// S.len, transform q reading src[q*st_col + i*st_el] and writing dst[q*dt_col + k*dt_el],
// optionally twiddled by W_Ltot^{q k}, processed in tiles of S.C columns (lanes column-fastest).
// pre/post hooks: conj and chirp handling are applied by the caller's flags.
template <typename Real, int E>
__device__ void run_phase(const RaggedSub &S, const void *src_raw, void *dst_raw, int64_t ncol,
                          int64_t st_col, int64_t st_el, int64_t dt_col, int64_t dt_el, bool twiddle,
                          const RaggedClass &K, bool conj_in, bool conj_out, double scale, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const V *src = reinterpret_cast<const V *>(src_raw);
    V *dst = reinterpret_cast<V *>(dst_raw);
    const int len = S.pt.n, T = len / E, C = S.C;
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const bool lane_ok = t < T;
    const V *tw = reinterpret_cast<const V *>(S.tw);
    const V *hi = reinterpret_cast<const V *>(K.hi);
    const V *lo = reinterpret_cast<const V *>(K.lo);
    const Real sc = (Real)scale;
    for (int64_t q0 = 0; q0 < ncol; q0 += C) {
        const int64_t q = q0 + c;
        const bool active = lane_ok && q < ncol;
        V reg[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            V v = active ? src[q * st_col + (int64_t)(t + i * T) * st_el] : c_make<V>(0, 0);
            reg[i] = conj_in ? c_conj(v) : v;
        }
        fft_core<E>(reg, lds + c, C, S.pt, tw, lane_ok ? t : 0, T, lane_ok);
        if (active) {
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const int k = t + i * T;
                V v = reg[i];
                if (twiddle) {
                    const int64_t e = (q * (int64_t)k) % K.Lx;
                    v = c_mul(v, c_mul(hi[e >> K.shift], lo[e & ((int64_t(1) << K.shift) - 1)]));
                }
                if (conj_out) v = c_conj(v);
                dst[q * dt_col + (int64_t)k * dt_el] = c_scale(v, sc);
            }
        }
        __syncthreads(); // the LDS tile is reused by the next columns
    }
}

template <typename Real, int EMAX>
__device__ void phase_dispatch(const RaggedSub &S, const void *src, void *dst, int64_t ncol,
                               int64_t st_col, int64_t st_el, int64_t dt_col, int64_t dt_el,
                               bool twiddle, const RaggedClass &K, bool cin, bool cout, double scale,
                               void *lds)
{
    switch (S.E) {
#define AFFT_RP(EV)                                                                              \
    case EV:                                                                                     \
        if constexpr (EV <= EMAX)                                                                \
            run_phase<Real, EV>(S, src, dst, ncol, st_col, st_el, dt_col, dt_el, twiddle, K, cin, \
                                cout, scale, lds);                                               \
        break;
        AFFT_RP(1) AFFT_RP(2) AFFT_RP(3) AFFT_RP(4) AFFT_RP(5) AFFT_RP(6) AFFT_RP(8) AFFT_RP(9)
        AFFT_RP(10) AFFT_RP(12) AFFT_RP(15) AFFT_RP(16)
#undef AFFT_RP
    default: break;
    }
}

// Four-step transform of length Lx = L1*L2 from src (element i at src[i]) into dst (natural
// order), through scratch `mid`: both phases by this workgroup.
template <typename Real, int EMAX>
__device__ void four_step_wg(const RaggedClass &K, const void *src, void *mid, void *dst, bool cin,
                             bool cout, double scale, void *lds)
{
    const int64_t L1 = K.s1.pt.n, L2 = K.s2.pt.n;
    // phase A: column n2 of the L1 x L2 view, FFT over n1, twiddle W^{n2 k1}, to mid[n2*L1 + k1]
    phase_dispatch<Real, EMAX>(K.s1, src, mid, L2, 1, L2, L1, 1, true, K, cin, false, 1.0, lds);
    phase_sync();
    // phase B: column k1 of mid (L2 x L1 view), FFT over n2, to dst[k1 + L1*k2]
    phase_dispatch<Real, EMAX>(K.s2, mid, dst, L1, 1, L1, 1, L1, false, K, false, cout, scale, lds);
    phase_sync();
}

template <typename Real, int EMAX>
__device__ void run_two_phase(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it,
                              void *lds, int64_t wg_slot)
{
    using V = typename cplx_of<Real>::type;
    const int64_t off = a.offsets[a.members[it.first]];
    const V *in = reinterpret_cast<const V *>(a.in) + off;
    V *out = reinterpret_cast<V *>(a.out) + off;
    V *scrA = reinterpret_cast<V *>(a.scratch) + wg_slot * 2 * a.scratch_elems;
    V *scrB = scrA + a.scratch_elems;
    const double inv = a.inverse ? 1.0 / (double)K.L : 1.0;
    if (K.kind == kTwoPhase) {
        four_step_wg<Real, EMAX>(K, in, scrA, out, a.conj, a.conj, inv, lds);
        return;
    }
    // chirp-z with two four-step FFT_M: x*c -> A, F(A) -> A (via B), conj(A*filt) -> A,
    // F(A) -> A (via B), y = conj(A) c
    const V *chirp = reinterpret_cast<const V *>(K.chirp);
    const V *filt = reinterpret_cast<const V *>(K.filt);
    const int64_t n = K.L, M = K.Lx;
    for (int64_t m = threadIdx.x; m < M; m += blockDim.x) {
        V v = c_make<V>(0, 0);
        if (m < n) {
            v = in[m];
            if (a.conj) v = c_conj(v);
            v = c_mul(v, chirp[m]);
        }
        scrA[m] = v;
    }
    phase_sync();
    four_step_wg<Real, EMAX>(K, scrA, scrB, scrA, false, false, 1.0, lds);
    for (int64_t m = threadIdx.x; m < M; m += blockDim.x) scrA[m] = c_conj(c_mul(scrA[m], filt[m]));
    phase_sync();
    four_step_wg<Real, EMAX>(K, scrA, scrB, scrA, false, false, 1.0, lds);
    const Real sc = (Real)inv;
    for (int64_t m = threadIdx.x; m < n; m += blockDim.x) {
        V v = c_mul(c_conj(scrA[m]), chirp[m]);
        if (a.conj) v = c_conj(v);
        out[m] = c_scale(v, sc);
    }
    phase_sync();
}

template <typename Real, int EMAX>
__global__ void __launch_bounds__(kRaggedBlock) afft_ragged(const RaggedArgs a)
{
    extern __shared__ __align__(16) unsigned char afft_smem[];
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = (int)atomicAdd(&a.queue[0], 1u);
        __syncthreads();
        const int idx = s_item;
        __syncthreads();
        if (idx >= a.nitems) break;
        const RaggedItem it = a.items[idx];
        const RaggedClass &K = a.classes[it.cls];
        if (K.kind == kDirect) {
            switch (K.E) {
#define AFFT_RD(EV)                                                                              \
    case EV:                                                                                     \
        if constexpr (EV <= EMAX) run_direct<Real, EV>(a, K, it, afft_smem);                     \
        break;
                AFFT_RD(1) AFFT_RD(2) AFFT_RD(3) AFFT_RD(4) AFFT_RD(5) AFFT_RD(6) AFFT_RD(8)
                AFFT_RD(9) AFFT_RD(10) AFFT_RD(12) AFFT_RD(15) AFFT_RD(16)
#undef AFFT_RD
            default: break;
            }
        } else if (K.kind == kBlue) {
            switch (K.E) {
#define AFFT_RB(EV)                                                                              \
    case EV:                                                                                     \
        if constexpr (EV <= EMAX) run_blue<Real, EV>(a, K, it, afft_smem);                       \
        break;
                AFFT_RB(1) AFFT_RB(2) AFFT_RB(3) AFFT_RB(4) AFFT_RB(5) AFFT_RB(6) AFFT_RB(8)
                AFFT_RB(9) AFFT_RB(10) AFFT_RB(12) AFFT_RB(15) AFFT_RB(16)
#undef AFFT_RB
            default: break;
            }
        } else {
            run_two_phase<Real, EMAX>(a, K, it, afft_smem, blockIdx.x);
        }
        __syncthreads(); // LDS reuse by the next item
    }
    // the last workgroup out resets the ticket counters for the next launch or graph replay
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned done = atomicAdd(&a.queue[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&a.queue[0], 0u);
            atomicExch(&a.queue[1], 0u);
        }
    }
}

hipError_t launch_ragged(int dtype, const RaggedArgs &a, int64_t grid, size_t lds, hipStream_t s)
{
    const void *fn = dtype == 0 ? reinterpret_cast<const void *>(afft_ragged<float, 16>)
                                : reinterpret_cast<const void *>(afft_ragged<double, 10>);
    if (lds > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
    }
    if (dtype == 0)
        hipLaunchKernelGGL((afft_ragged<float, 16>), dim3((unsigned)grid), dim3(kRaggedBlock), lds, s, a);
    else
        hipLaunchKernelGGL((afft_ragged<double, 10>), dim3((unsigned)grid), dim3(kRaggedBlock), lds, s, a);
    return hipGetLastError();
}

int ragged_occupancy(int dtype, size_t lds)
{
    const void *fn = dtype == 0 ? reinterpret_cast<const void *>(afft_ragged<float, 16>)
                                : reinterpret_cast<const void *>(afft_ragged<double, 10>);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, kRaggedBlock, lds) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

} // namespace afft
