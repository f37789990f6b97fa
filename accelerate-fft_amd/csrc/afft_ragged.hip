// afft_ragged.hip -- ragged batches: many 1D transforms of mixed lengths, stored back to back,
// in one persistent launch (kernel afft_ragged<Real, EMAX>).
//
// The host sorts the transforms into length classes and cuts the classes that fit one workgroup
// into work items, ordered by decreasing cost.  The workgroups walk the device-side item list
// round-robin (a grid of up to four times the resident workgroups: one item each for all but the
// largest batches, so a workgroup that has issued its stores retires and the next one's loads
// overlap them); no workgroup ever waits for another, so the launch always drains, and there is
// no device-side state to reset between launches: the launch replays as a captured hipGraph.  Longer classes run as plans of the batched engine over the rows of the class
// (gathered by offset), in the same stream and graph.
//
// Item kinds:
//   kDirect     C transforms of length L <= one workgroup's layout, Stockham passes in LDS
//   kBlue       C transforms of a non-5-smooth length with padded length M <= one workgroup:
//               the fused chirp-z transform (two forward passes sharing one LDS slot)
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
    // transform c's threads are [c*T, c*T + T): within one wavefront when T divides 64
    const bool wave = T <= 64 && (64 % T) == 0;
    fft_core<E>(reg, slot, 1, K.pt, reinterpret_cast<const V *>(K.tw), t, T, c < K.C, wave);
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
    const bool wave = T <= 64 && (64 % T) == 0;
    fft_core<E>(reg, slot, 1, K.pt, tw, t, T, on, wave);
#pragma unroll
    for (int i = 0; i < E; ++i) reg[i] = c_conj(c_mul(reg[i], on ? filt[t + i * T] : c_make<V>(0, 0)));
    // the second transform's first exchange writes the slot the first one read last: in order
    // within the wavefront, or behind fft_core's leading barrier
    fft_core<E>(reg, slot, 1, K.pt, tw, t, T, on, wave);
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

// A batch of C = kRaggedBlock / T transforms of a compiled layout (direct, or the two transforms
// of a chirp-z class of padded length N): the passes of CtStockham, every radix, stride and
// twiddle index a compile-time constant.  Transforms of up to 64 threads exchange within their
// wavefront.
template <typename Real, int N, int E, int... Rs>
__device__ __noinline__ void run_ct(const RaggedArgs &a, const RaggedClass &K, const RaggedItem &it, void *lds_raw)
{
    using V = typename cplx_of<Real>::type;
    using CT = CtStockham<V, N, E, Rs...>;
    constexpr int T = N / E;
    constexpr bool kWave = T <= 64 && 64 % T == 0;
    static_assert(kRaggedBlock % T == 0, "whole transforms per workgroup");
    V *lds = reinterpret_cast<V *>(lds_raw);
    const int t = threadIdx.x % T, c = threadIdx.x / T;
    const bool active = c < it.count;
    const bool blue = K.kind == kBlue;
    const int n = K.L; // the transform's own length: N, or the chirp-z length below N
    const int64_t off = active ? a.moffs[it.first + c] : 0;
    // global-memory pointers: the item's stores then count only against the vector-memory
    // counter, so neither the LDS waits of the next item nor the barrier ending this one wait for
    // them (through the generic pointers of a non-inlined function they would be flat stores)
    GIn<V> *in = (GIn<V> *)a.in;
    GOut<V> *out = (GOut<V> *)a.out;
    GIn<V> *chirp = (GIn<V> *)K.chirp;
    CT ct;
    ct.load((GIn<V> *)K.tw, t);
    V reg[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int m = t + i * T;
        V v = c_make<V>(0, 0);
        if (active && m < n) {
            v = ld_elem(in + off + m);
            if (a.conj) v = c_conj(v);
            if (blue) v = c_mul(v, ld_elem(chirp + m));
        }
        reg[i] = v;
    }
    V *slot = lds + c * lds_padded_len(N);
    int tl = t;
    __asm__ volatile("" : "+v"(tl));
    ct.template run<true, kWave, true>(reg, slot, 1, tl);
    if (blue) {
        GIn<V> *filt = (GIn<V> *)K.filt;
#pragma unroll
        for (int i = 0; i < E; ++i) reg[i] = c_conj(c_mul(reg[i], ld_elem(filt + t + i * T)));
        // the second transform's first exchange writes the slot the first one read last: in
        // order within the wavefront, or behind the pass's leading barrier
        ct.template run<true, kWave, true>(reg, slot, 1, tl);
    }
    const Real sc = a.inverse ? (Real)(1.0 / (double)n) : (Real)1;
    if (active) {
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int m = t + i * T;
            if (m < n) {
                V v = blue ? c_mul(c_conj(reg[i]), ld_elem(chirp + m)) : reg[i];
                if (a.conj) v = c_conj(v);
                st_elem(out + off + m, c_scale(v, sc));
            }
        }
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
            if constexpr (sizeof(Real) == 4) {
                if (K.ctn) {
                    switch (K.ctn) {
#define AFFT_RCT(N, E, ...)                                                                      \
    case N: run_ct<Real, N, E, __VA_ARGS__>(a, K, it, afft_smem); break;
                        AFFT_RAGGED_CT_LAYOUTS(AFFT_RCT)
#undef AFFT_RCT
                    default: break;
                    }
                    lds_sync(); // LDS (transforms and descriptor) reuse by the next item
                    continue;
                }
            }
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
            } else {
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
            }
            __syncthreads(); // LDS (transforms and descriptor) reuse by the next item
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
