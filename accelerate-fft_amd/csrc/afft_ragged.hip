// afft_ragged.hip -- ragged batches: many 1D transforms of mixed lengths, stored back to back,
// in one persistent launch per phase (kernel afft_ragged<Real, EMAX>).
//
// The host sorts the transforms into length classes and cuts them into work items, ordered by
// decreasing cost within each phase.  The resident workgroups walk the phase's device-side item
// list round-robin; no workgroup ever waits for another, so every launch drains, and there is no
// device-side state to reset between launches: the launches replay as one captured hipGraph.
//
// Item kinds:
//   kDirect     C transforms of length L <= one workgroup's layout, Stockham passes in LDS
//   kBlue       C transforms of a non-5-smooth length with padded length M <= one workgroup:
//               the fused chirp-z transform (two forward passes sharing one LDS slot)
//   kTwoPhase   one column tile of one stage of a long 5-smooth transform (four-step, L = L1 * L2):
//               stage A tiles write the twiddled transposed columns into the output, stage B
//               tiles transform them in place
//   kBlueTwo    one column tile of one of the four stages of a long chirp-z transform, through
//               two scratch buffers
// Stage s of every long transform runs in phase s, a launch of its own: the launch boundary
// makes a stage's stores visible to every XCD, which costs far less than device-scope counters
// or grid barriers inside one launch, and the tiles of a stage spread over the whole grid.
#include <hip/hip_runtime.h>

#include "afft_device.hpp"
#include "afft_ragged.hpp"

namespace afft {

// A batch of C same-length transforms at arbitrary element offsets.
template <typename Real, int E>
__device__ __forceinline__ void run_direct(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const int L = K.L, T = L / E;
    const int t = threadIdx.x % T, c = threadIdx.x / T;
    const bool active = c < it.count;
    const int64_t off = active ? a.moffs[it.first + c] : 0;
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
__device__ __forceinline__ void run_blue(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const int n = K.L, M = K.M, T = M / E;
    const int t = threadIdx.x % T, c = threadIdx.x / T;
    const bool active = c < it.count;
    const int64_t off = active ? a.moffs[it.first + c] : 0;
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

// One column tile of one stage of a long transform (Lx = L1 * L2, four-step scheme).
//   A stages: column q < L2 of the L1 x L2 view, element n1 at src[n1*L2 + q]; output k1 goes
//             to dst[q*L1 + k1] multiplied by W_Lx^{q k1} (q k1 < Lx: no reduction needed)
//   B stages: column q < L1, element n2 at src[n2*L1 + q]; output k2 goes to dst[q + L1*k2]
// Four-step Bluestein (M = Lx): stage 0 multiplies the input by the chirp on load (zero beyond
// n), stage 1 applies conj(. * filter) on store, stage 3 applies conj(.) * chirp on store.
template <typename Real, int E>
__device__ __forceinline__ void run_tile(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    const int stage = it.stage, tile = it.tile;
    using V = typename cplx_of<Real>::type;
    V *lds = reinterpret_cast<V *>(lds_raw);
    const bool blue = K.kind == kBlueTwo;
    const bool isA = (stage % 2) == 0;
    const RaggedSub &S = isA ? K.s1 : K.s2;
    const int64_t L1 = K.s1.pt.n, Lx = K.Lx;
    const int64_t n = K.L;
    const int64_t off = a.moffs[it.first];
    const V *in = reinterpret_cast<const V *>(a.in) + off;
    V *out = reinterpret_cast<V *>(a.out) + off;
    // four-step: A in -> out (transposed, twiddled), B out -> out in place (a B tile reads and
    // writes the same column set).  Bluestein: A1 in -> scrA, B1 scrA in place, A2 scrA -> scrB,
    // B2 scrB -> out.
    V *scrA = blue ? reinterpret_cast<V *>(a.scratch) + a.big_off[it.big] : nullptr;
    V *scrB = blue ? scrA + Lx : nullptr;
    const V *src;
    V *dst;
    if (!blue) {
        src = isA ? in : out;
        dst = out;
    } else {
        src = stage == 0 ? in : (stage == 1 ? scrA : (stage == 2 ? scrA : scrB));
        dst = stage == 0 ? scrA : (stage == 1 ? scrA : (stage == 2 ? scrB : out));
    }
    const int64_t L2 = Lx / L1;
    const int64_t ncol = isA ? L2 : L1;
    const int64_t st_el = isA ? L2 : L1;
    const int64_t dt_col = isA ? L1 : 1, dt_el = isA ? 1 : L1;
    const int len = S.pt.n, T = len / E, C = S.C;
    const int c = threadIdx.x % C, t = threadIdx.x / C;
    const bool lane_ok = t < T;
    const int64_t q = (int64_t)tile * C + c;
    const bool active = lane_ok && q < ncol;
    const V *chirp = reinterpret_cast<const V *>(K.chirp);
    const V *filt = reinterpret_cast<const V *>(K.filt);
    const bool first = stage == 0, last = stage == (blue ? kStagesBlue : kStagesTwo) - 1;
    V reg[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        V v = c_make<V>(0, 0);
        if (active) {
            const int64_t pos = (int64_t)(t + i * T) * st_el + q;
            if (blue && first) {
                if (pos < n) {
                    v = src[pos];
                    if (a.conj) v = c_conj(v);
                    v = c_mul(v, chirp[pos]);
                }
            } else {
                v = src[pos];
                if (first && a.conj) v = c_conj(v);
            }
        }
        reg[i] = v;
    }
    fft_core<E>(reg, lds + c, C, S.pt, reinterpret_cast<const V *>(S.tw), lane_ok ? t : 0, T, lane_ok);
    if (active) {
        const V *hi = reinterpret_cast<const V *>(K.hi);
        const V *lo = reinterpret_cast<const V *>(K.lo);
        const Real sc = a.inverse ? (Real)(1.0 / (double)n) : (Real)1;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int64_t k = t + i * T;
            const int64_t dpos = q * dt_col + k * dt_el;
            V v = reg[i];
            if (isA) {
                const int64_t e = q * k;
                v = c_mul(v, c_mul(hi[e >> K.shift], lo[e & ((int64_t(1) << K.shift) - 1)]));
                dst[dpos] = v;
            } else if (!blue) {
                if (a.conj) v = c_conj(v);
                dst[dpos] = c_scale(v, sc);
            } else if (stage == 1) {
                dst[dpos] = c_conj(c_mul(v, filt[dpos]));
            } else if (dpos < n) {
                v = c_mul(c_conj(v), chirp[dpos]);
                if (a.conj) v = c_conj(v);
                dst[dpos] = c_scale(v, sc);
            }
        }
    }
    (void)last;
}

template <typename Real, int EMAX>
__device__ __forceinline__ void tile_dispatch(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds)
{
    const RaggedSub &S = (it.stage % 2 == 0) ? K.s1 : K.s2;
    switch (S.E) {
#define AFFT_RT(EV)                                                                              \
    case EV:                                                                                     \
        if constexpr (EV <= EMAX) run_tile<Real, EV>(a, K, it, lds);                             \
        break;
        AFFT_RT(1) AFFT_RT(2) AFFT_RT(3) AFFT_RT(4) AFFT_RT(5) AFFT_RT(6) AFFT_RT(8) AFFT_RT(9)
        AFFT_RT(10) AFFT_RT(12) AFFT_RT(15) AFFT_RT(16)
#undef AFFT_RT
    default: break;
    }
}

template <typename Real, int EMAX>
__global__ void __launch_bounds__(kRaggedBlock) __attribute__((amdgpu_waves_per_eu(3, 8))) afft_ragged(const RaggedArgs a)
{
    extern __shared__ __align__(16) unsigned char afft_smem[];
    // the current item's class descriptor: its pass tables are read on every pass, so they are
    // staged in LDS once per item instead of costing a global-memory round trip per pass
    __shared__ __align__(16) RaggedClass sK;
    static_assert(sizeof(RaggedClass) % 4 == 0, "descriptor copied as dwords");
    {
        // items are sorted by decreasing cost; workgroup w takes w, w + G, w + 2G ... (a shared
        // ticket counter serialises on one address across the XCDs and costs more than the
        // balance it buys)
        for (int idx = blockIdx.x; idx < a.nitems; idx += gridDim.x) {
            const RaggedItem it = a.items[idx];
            {
                const uint32_t *src = reinterpret_cast<const uint32_t *>(a.classes + it.cls);
                uint32_t *dst = reinterpret_cast<uint32_t *>(&sK);
                for (int w = threadIdx.x; w < (int)(sizeof(RaggedClass) / 4); w += blockDim.x) dst[w] = src[w];
            }
            __syncthreads();
            const RaggedClass &K = sK;
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
                // count consecutive column tiles of one stage of a long transform
                RaggedItem t = it;
                for (int k = 0; k < it.count; ++k) {
                    t.tile = it.tile + k;
                    tile_dispatch<Real, EMAX>(a, K, t, afft_smem);
                    __syncthreads();
                }
            }
            __syncthreads(); // LDS (tiles and descriptor) reuse by the next item
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
