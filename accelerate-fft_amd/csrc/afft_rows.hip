// afft_rows.hip -- persistent, software-pipelined kernels for batches of contiguous rows of the
// common lengths (the `fft` innermost-axis hot path, FFT.hs:63-84).
//
// Kernel names: afft_rows<Real, N, E, PF, W, R0, R1, ...>  (N = length, E = elements per thread,
// PF = prefetch the next row, W = waves per SIMD the register allocation must allow, R = pass
// radices).
//
// Why a second kernel family next to afft_stockham: the batched transform is HBM-bound, and the
// generic kernel leaves the memory system idle while a workgroup computes (LDS caps a CU at a
// few resident rows).  Here every workgroup loops over row tiles and issues the loads of tile
// g + gridDim.x before computing tile g, so each resident workgroup keeps a second row in
// flight in registers.  Everything the generic kernel computes at run time (pass strides,
// modulo, addresses) is a compile-time constant here.
//
// Ordering constraint of the hardware: vector-memory loads and stores retire in issue order
// (one vmcnt counter), so a wait for any load also waits for every older load and store.  The
// row-invariant twiddle bases are therefore loaded once before the row loop; inside the loop
// the only vector-memory operations are the prefetch and the stores, and the LDS exchanges
// synchronise with LDS-only fences (s_waitcnt lgkmcnt + s_barrier), which leave the prefetch
// and the stores in flight.
//
// Twiddles: pass p (radix R, stride Ns) needs W_{Ns R}^{r k}, r < R, k = b mod Ns.  Each
// thread loads W^k and W^{4k} from the FP64-generated table W_N^m (rounded to Real) and forms
// the others by at most three complex products (W^2k = (W^k)^2, W^3k = W^2k W^k, W^(4a+b)k =
// W^4ak W^bk ...), i.e. within a few ulp of the correctly rounded value.
#include <hip/hip_runtime.h>

#include <utility>

#include <algorithm>
#include <cstdlib>

#include "afft_device.hpp"
#include "afft_internal.hpp"

namespace afft {

// Synchronisation of an LDS exchange among the T threads of one row.  When they all sit in one
// wavefront (T divides 64; a row's threads are contiguous), the exchange is private to that
// wavefront, whose LDS instructions execute in order: ordering the accesses at wavefront scope
// is enough, and no workgroup barrier stalls the other rows of the workgroup.
template <int T> __device__ __forceinline__ void row_sync()
{
    if constexpr (T <= 64 && 64 % T == 0) wave_sync();
    else lds_sync();
}

// T threads per row, C rows per tile (a workgroup of about 256 threads), at least CM rows: the
// transposed-store form needs 128-byte runs of neighbouring rows (16 in FP32), or the runs of
// neighbouring tiles share lines and the partial lines cost extra write traffic (measured on the
// 2^22 split six-step with 8-row tiles: 1.48x the array written)
template <int N, int E, int CM = 1> struct RowShape {
    static constexpr int T = N / E;
    static constexpr int C0 = T >= 256 ? 1 : 256 / T;
    static constexpr int C = C0 >= CM || CM * T > 1024 ? C0 : CM;
    static constexpr int BLOCK = C * T;
};

// minimum rows per tile of the transposed-store form: one 128-byte line of elements
template <typename Real, bool TR> constexpr int row_min_tile() { return TR ? 64 / (int)sizeof(Real) : 1; }

// Position of element p of a row in the one-row array of reals of the split exchange: the low
// five bits permuted by the next five, so that the stride-R scatters of the early passes (R
// consecutive lanes' elements R apart) spread over the banks without padding; a bijection on
// every aligned block of 32, so the array is exactly N reals long.
__device__ __forceinline__ int lds_swz(int p) { return p ^ ((p >> 5) & 31); }

// SPLIT: rows too long for a complex LDS slot (a row of 20480 points is 160 KB of Complex Float)
// exchange through one array of N reals, the real parts and then the imaginary parts; one row
// per workgroup, no prefetch (its registers hold the row).
// RES: the rows loaded are alpha (in - in2), the difference of two arrays (the residual of a
// squared-error loss, transformed without being stored first).
template <typename Real, int N, int E, bool PF, int W, bool LD, bool TR, bool SPLIT, bool RES, int... Rs>
__global__ void __launch_bounds__((RowShape<N, E, row_min_tile<Real, TR>()>::BLOCK))
__attribute__((amdgpu_waves_per_eu(W, 8)))
afft_rows(const RowArgs a)
{
    using V = typename cplx_of<Real>::type;
    using RL = Radices<Rs...>;
    using Shape = RowShape<N, E, row_min_tile<Real, TR>()>;
    constexpr int T = Shape::T;
    constexpr int C = Shape::C;
    constexpr int NP = RL::NP;
    constexpr int PADN = lds_padded_len(N);
    static_assert(T * E == N, "E must divide N");
    static_assert(!SPLIT || (C == 1 && !TR && !PF && N % 32 == 0), "split exchange: one row, no prefetch");

    extern __shared__ __align__(16) unsigned char afft_smem[];
    V *lds = reinterpret_cast<V *>(afft_smem);
    const int t = threadIdx.x % T;
    const int c = threadIdx.x / T;
    V *slot = lds + c * PADN;
    const V *__restrict__ in = reinterpret_cast<const V *>(a.in);
    V *__restrict__ out = reinterpret_cast<V *>(a.out);
    const V *__restrict__ tw = reinterpret_cast<const V *>(a.tw);
    const int64_t nrows = a.nrows;
    const int64_t ntiles = (nrows + C - 1) / C;
    const Real sc = (Real)a.scale;
    const bool cin = a.conj_in, cout = a.conj_out;

    const bool nt_in = a.nt_in, nt_out = a.nt_out;
    const int64_t *__restrict__ roffs = a.roffs;
    // transposed stores of transforms that start anywhere (tr_rot): tile slot r of transform o
    // holds row k1 = (u - s) mod n1, u = r - o n1, s = address of X_o[0] in elements modulo C.
    // The tile's C rows then store runs base - s + jC + [0, C) at every k2 (the rows past
    // n1 - s in the run of k2 + 1), which are whole aligned lines written by this workgroup
    // alone, where unrotated tiles would write runs straddling the lines of the next tile.
    auto tr_row = [&](int64_t row, int64_t &base, int64_t &k1) -> int64_t {
        const int64_t n1 = a.tr_n1;
        const int64_t o = row / n1;
        const int64_t u = row - o * n1;
        base = a.tr_toff ? a.tr_toff[o] : o * (int64_t)N * n1;
        k1 = u;
        if (a.tr_rot) {
            const int64_t s = (int64_t)((reinterpret_cast<uintptr_t>(out) / sizeof(V) + (uint64_t)base) % C);
            k1 = u >= s ? u - s : u - s + n1;
        }
        return o * n1 + k1;
    };
    // the row a tile slot reads: itself, or its rotated row
    auto src_row = [&](int64_t row) -> int64_t {
        if constexpr (TR) {
            if (a.tr_rot) {
                int64_t base, k1;
                return tr_row(row, base, k1);
            }
        }
        return row;
    };
    V nxt[E];
    V nxt2[RES ? E : 1];
    const V *__restrict__ in2 = reinterpret_cast<const V *>(a.in2);
    auto load_tile = [&](int64_t tile) {
        if (tile < ntiles && tile * C + c < nrows) {
            const int64_t row = src_row(tile * C + c);
            const V *p = in + (roffs ? roffs[row] : row * N) + t;
            if (nt_in) {
#pragma unroll
                for (int i = 0; i < E; ++i) nxt[i] = c_load_nt(p + i * T);
            } else {
#pragma unroll
                for (int i = 0; i < E; ++i) nxt[i] = p[i * T];
            }
            if constexpr (RES) {
                const V *q = in2 + (roffs ? roffs[row] : row * N) + t;
#pragma unroll
                for (int i = 0; i < E; ++i) nxt2[i] = nt_in ? c_load_nt(q + i * T) : q[i * T];
            }
        }
    };

    // base twiddles of every pass: they depend on the thread, not on the row, so they are
    // loaded once, before any row load, and stay in registers
    V tb[NP > 0 ? NP : 1][2 * E];
    static_for<NP>([&](auto pc) {
        constexpr int p = decltype(pc)::value;
        constexpr int R = RL::R[p];
        constexpr int Ns = RL::ns(p);
        if constexpr (Ns > 1) {
            constexpr int B = E / R;
            // the last pass (Ns * R == N) has k = t + j*T exactly: one base W^t per thread, the
            // factor W^{jT} of butterfly j is a compile-time constant
            constexpr int NB = (T % Ns == 0 || Ns * R == N) ? 1 : B;
            constexpr int STEP = N / (Ns * R);
#pragma unroll
            for (int jb = 0; jb < NB; ++jb) {
                const int k = (t + jb * T) % Ns;
                tb[p][2 * jb] = tw[k * STEP];
                tb[p][2 * jb + 1] = (R > 5) ? tw[4 * k * STEP] : tb[p][2 * jb];
            }
        }
    });

    // work item g -> tile.  Transposed stores map the items XCD-contiguously (XCD x, which gets
    // the workgroups b = x mod 8, takes one contiguous range of tiles): neighbouring tiles write
    // the two halves of the same output lines, and on one XCD those halves meet in one L2
    auto tile_of = [&](int64_t item) -> int64_t { return TR ? xcd_tile(item, ntiles) : item; };
    int64_t g = blockIdx.x;
    if constexpr (PF) load_tile(tile_of(g));
    for (; g < ntiles; g += gridDim.x) {
        const int64_t gt = tile_of(g);
        if constexpr (!PF) load_tile(gt);
        // LD: the LDS addresses depend only on the thread; passing its index through an empty
        // asm once per row makes them recomputed here instead of hoisted out of the loop, where
        // with many passes (or a tight register budget) they would keep many registers live
        int tl = t;
        if constexpr (LD) __asm__ volatile("" : "+v"(tl));
        V reg[E];
        if constexpr (RES) {
            const Real al = (Real)a.alpha;
#pragma unroll
            for (int i = 0; i < E; ++i) {
                const V v = c_make<V>(al * (nxt[i].x - nxt2[i].x), al * (nxt[i].y - nxt2[i].y));
                reg[i] = cin ? c_conj(v) : v;
            }
        } else {
            // (this exact form: written through a temporary, the compiler merged the streaming
            // and the plain load paths and dropped the streaming hint, 7% on the 4096-point rows)
#pragma unroll
            for (int i = 0; i < E; ++i) reg[i] = cin ? c_conj(nxt[i]) : nxt[i];
        }
        if (a.sg_in.on && gt * C + c < nrows) {
            const int64_t r0 = src_row(gt * C + c);
            const int64_t rb = roffs ? roffs[r0] : r0 * N;
#pragma unroll
            for (int i = 0; i < E; ++i) reg[i] = sign_apply(a.sg_in, rb + t + i * T, reg[i]);
        }
        if constexpr (PF) {
            __builtin_amdgcn_sched_barrier(0);
            load_tile(tile_of(g + gridDim.x));
            __builtin_amdgcn_sched_barrier(0);
        }

        static_for<NP>([&](auto pc) {
            constexpr int p = decltype(pc)::value;
            constexpr int R = RL::R[p];
            constexpr int Ns = RL::ns(p);
            constexpr int B = E / R;
            static_for<B>([&](auto jc) {
                constexpr int j = decltype(jc)::value;
                V v[R];
#pragma unroll
                for (int r = 0; r < R; ++r) v[r] = reg[j + r * B];
                if constexpr (Ns > 1) {
                    if constexpr (T % Ns == 0) {
                        apply_twiddles<R>(v, tb[p][0], tb[p][1]);
                    } else if constexpr (Ns * R == N) {
                        const V b1 = twiddle_const<N, j * T>(tb[p][0]);
                        const V b4 = R > 5 ? twiddle_const<N, 4 * j * T>(tb[p][1]) : b1;
                        apply_twiddles<R>(v, b1, b4);
                    } else {
                        apply_twiddles<R>(v, tb[p][2 * j], tb[p][2 * j + 1]);
                    }
                }
                dft<R>::run(v);
#pragma unroll
                for (int r = 0; r < R; ++r) reg[j + r * B] = v[r];
            });
            if constexpr (p + 1 < NP && SPLIT) {
                Real *fl = reinterpret_cast<Real *>(afft_smem);
                // real parts, then imaginary parts: the second round scatters the .y halves,
                // which the first round's reads (into .x) left as they were.  The positions are
                // recomputed per round from the thread index (through an empty asm): kept live
                // across the passes they would not fit the register budget of 1024 threads.
                static_for<2>([&](auto hc) {
                    constexpr int h = decltype(hc)::value;
                    int tp = tl;
                    __asm__ volatile("" : "+v"(tp));
                    lds_sync(); // the array's previous contents have been read
#pragma unroll
                    for (int j = 0; j < B; ++j) {
                        const int b = tp + j * T;
                        const int k = b % Ns;
                        const int dst = (b - k) * R + k;
#pragma unroll
                        for (int r = 0; r < R; ++r) fl[lds_swz(dst + r * Ns)] = h ? reg[j + r * B].y : reg[j + r * B].x;
                    }
                    lds_sync();
#pragma unroll
                    for (int i = 0; i < E; ++i) {
                        const Real v = fl[lds_swz(tp + i * T)];
                        if constexpr (h) reg[i].y = v;
                        else reg[i].x = v;
                    }
                });
            } else if constexpr (p + 1 < NP) {
                row_sync<T>(); // the previous exchange has been read by every thread of the row
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    const int b = tl + j * T;
                    const int k = b % Ns;
                    const int dst = (b - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) slot[lds_pad(dst + r * Ns)] = reg[j + r * B];
                }
                row_sync<T>();
#pragma unroll
                for (int i = 0; i < E; ++i) reg[i] = slot[lds_pad(tl + i * T)];
            }
        });

        if constexpr (TR) {
            // transposed store (the last pass of a six-step transform): the C rows of the tile
            // are consecutive k1 of one transform, and X[k1 + n1 k2] goes to k1 + n1 k2.  The
            // results are regrouped through LDS (element k2 of row c at k2 (C+1) + c) and re-read
            // with the row varying fastest across lanes, so each store writes C neighbouring k1.
            constexpr int C1 = C + 1;
            lds_sync(); // every row of the tile is past its last exchange read
#pragma unroll
            for (int i = 0; i < E; ++i) lds[(tl + i * T) * C1 + c] = reg[i];
            lds_sync();
            const int c2 = threadIdx.x % C, t2 = threadIdx.x / C;
            const int64_t row2 = gt * C + c2;
            if (row2 < nrows) {
                const int64_t n1 = a.tr_n1;
                int64_t base, k1;
                (void)tr_row(row2, base, k1);
                const int64_t ob = base + k1;
#pragma unroll
                for (int i = 0; i < E; ++i) {
                    const int64_t k2 = t2 + i * T;
                    V v = lds[k2 * C1 + c2];
                    v = c_scale(cout ? c_conj(v) : v, sc);
                    v = sign_apply(a.sg_out, ob + n1 * k2, v);
                    if (nt_out) c_store_nt(out + ob + n1 * k2, v);
                    else out[ob + n1 * k2] = v;
                }
            }
            lds_sync(); // the regroup buffer is read before the next tile's first exchange
        } else {
            const int64_t row = gt * C + c;
            if (row < nrows && a.sg_out.on) {
                const int64_t rb = roffs ? roffs[row] : row * N;
#pragma unroll
                for (int i = 0; i < E; ++i) reg[i] = sign_apply(a.sg_out, rb + t + i * T, reg[i]);
            }
            if (row < nrows) {
                V *q = out + (roffs ? roffs[row] : row * N) + t;
                if (nt_out) {
#pragma unroll
                    for (int i = 0; i < E; ++i) c_store_nt(q + i * T, c_scale(cout ? c_conj(reg[i]) : reg[i], sc));
                } else {
#pragma unroll
                    for (int i = 0; i < E; ++i) q[i * T] = c_scale(cout ? c_conj(reg[i]) : reg[i], sc);
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// Registry of the instantiated (dtype, N) layouts.

namespace {

using RowFn = void (*)(const RowArgs);

struct RowEntry {
    int dtype;
    int n;
    int E;
    bool pf;
    RowFn fn;
    int block;
    bool tr;
    int rows;   // rows per tile
    bool split; // exchanges through N reals (one row per workgroup)
    bool res;   // loads alpha (in - in2) (selected with variant kRowResidual)
};

#define AFFT_ROW(REAL, DT, N, E, PF, W, LD, ...)                                                 \
    RowEntry{DT, N, E, PF, afft_rows<REAL, N, E, PF, W, LD, false, false, false, __VA_ARGS__>,   \
             RowShape<N, E>::BLOCK, false, RowShape<N, E>::C, false, false}
// the transposed-store form, one per length (selected with variant kRowTransposed)
#define AFFT_ROW_TR(REAL, DT, N, E, PF, W, LD, ...)                                              \
    RowEntry{DT, N, E, PF, afft_rows<REAL, N, E, PF, W, LD, true, false, false, __VA_ARGS__>,    \
             RowShape<N, E, row_min_tile<REAL, true>()>::BLOCK, true, RowShape<N, E, row_min_tile<REAL, true>()>::C, \
             false, false}
// rows beyond a complex LDS slot: the split exchange
#define AFFT_ROW_SPLIT(REAL, DT, N, E, W, LD, ...)                                               \
    RowEntry{DT, N, E, false, afft_rows<REAL, N, E, false, W, LD, false, true, false, __VA_ARGS__>, \
             RowShape<N, E>::BLOCK, false, 1, true, false}
// the residual-load form, one per length (the first row pass of the squared-error spectral VJP)
#define AFFT_ROW_RES(REAL, DT, N, E, PF, W, LD, ...)                                             \
    RowEntry{DT, N, E, PF, afft_rows<REAL, N, E, PF, W, LD, false, false, true, __VA_ARGS__>,    \
             RowShape<N, E>::BLOCK, false, RowShape<N, E>::C, false, true}

// The first entry of a (dtype, n) is the default; the others are tuning alternatives.
const RowEntry kRows[] = {
    AFFT_ROW(float, 0, 256, 16, true, 2, false, 16, 16),
    AFFT_ROW(float, 0, 512, 16, true, 2, false, 16, 16, 2),
    AFFT_ROW(float, 0, 1024, 16, true, 2, false, 16, 16, 4),
    AFFT_ROW(float, 0, 1024, 16, false, 3, true, 16, 16, 4),
    AFFT_ROW(float, 0, 2048, 16, true, 2, false, 16, 16, 8),
    AFFT_ROW(float, 0, 2048, 16, false, 3, true, 16, 16, 8),
    AFFT_ROW(float, 0, 4096, 16, true, 2, false, 16, 16, 16),
    AFFT_ROW(float, 0, 4096, 16, true, 3, true, 16, 16, 16),
    AFFT_ROW(float, 0, 4096, 16, false, 3, true, 16, 16, 16),
    AFFT_ROW(float, 0, 8192, 16, true, 2, false, 16, 16, 16, 2),
    AFFT_ROW(float, 0, 8192, 16, false, 2, false, 16, 16, 16, 2),
    // mixed radices: the odd radix last, so that every earlier pass keeps one twiddle base
    AFFT_ROW(float, 0, 5120, 20, true, 2, true, 4, 4, 4, 4, 4, 5),
    AFFT_ROW(float, 0, 6144, 24, true, 2, true, 4, 8, 8, 8, 3),
    AFFT_ROW(float, 0, 10240, 20, true, 2, true, 2, 4, 4, 4, 4, 4, 5),
    AFFT_ROW(float, 0, 12288, 24, true, 2, true, 8, 8, 8, 8, 3),
    AFFT_ROW(float, 0, 12288, 24, false, 2, true, 8, 8, 8, 8, 3),
    // one row per workgroup through the split exchange: 16384 and 20480, whose rows fit the 128
    // registers of 1024 threads (24576, 32768 and 40960 did not: their spills made them slower
    // than two passes); every pass but the last has stride <= T
    AFFT_ROW_SPLIT(float, 0, 16384, 16, 4, true, 16, 16, 4, 16),
    AFFT_ROW_SPLIT(float, 0, 20480, 20, 4, true, 4, 4, 4, 4, 4, 4, 5),

    // the row lengths of six-step transforms n = 64 * n2
    AFFT_ROW(float, 0, 160, 20, true, 2, true, 2, 4, 4, 5),
    AFFT_ROW(float, 0, 320, 20, true, 2, true, 4, 4, 4, 5),
    AFFT_ROW(float, 0, 384, 24, true, 2, true, 4, 4, 8, 3),
    AFFT_ROW(float, 0, 640, 20, true, 2, true, 4, 4, 2, 4, 5),
    AFFT_ROW(float, 0, 768, 24, true, 2, true, 4, 8, 8, 3),
    AFFT_ROW(float, 0, 1536, 24, true, 2, true, 8, 8, 8, 3),
    AFFT_ROW_TR(float, 0, 160, 20, true, 2, true, 2, 4, 4, 5),
    AFFT_ROW_TR(float, 0, 320, 20, true, 2, true, 4, 4, 4, 5),
    AFFT_ROW_TR(float, 0, 384, 24, true, 2, true, 4, 4, 8, 3),
    AFFT_ROW_TR(float, 0, 512, 16, true, 2, true, 16, 16, 2),
    AFFT_ROW_TR(float, 0, 640, 20, true, 2, true, 4, 4, 2, 4, 5),
    AFFT_ROW_TR(float, 0, 768, 24, true, 2, true, 4, 8, 8, 3),
    AFFT_ROW(double, 1, 256, 8, true, 2, false, 8, 8, 4),
    AFFT_ROW(double, 1, 512, 8, true, 2, false, 8, 8, 8),
    AFFT_ROW(double, 1, 1024, 8, true, 2, false, 8, 8, 8, 2),
    AFFT_ROW(double, 1, 1024, 8, false, 2, false, 8, 8, 8, 2),
    AFFT_ROW(double, 1, 2048, 8, true, 2, false, 8, 8, 8, 4),
    AFFT_ROW(double, 1, 2048, 8, false, 2, false, 8, 8, 8, 4),
    AFFT_ROW(double, 1, 4096, 8, true, 2, false, 8, 8, 8, 8),
    AFFT_ROW(double, 1, 4096, 8, false, 2, false, 8, 8, 8, 8),
    // residual loads, with the default layouts of the image widths of spectral convolutions
    AFFT_ROW_RES(float, 0, 256, 16, true, 2, false, 16, 16),
    AFFT_ROW_RES(float, 0, 512, 16, true, 2, false, 16, 16, 2),
    AFFT_ROW_RES(float, 0, 1024, 16, true, 2, false, 16, 16, 4),
    AFFT_ROW_RES(double, 1, 256, 8, true, 2, false, 8, 8, 4),
    AFFT_ROW_RES(double, 1, 512, 8, true, 2, false, 8, 8, 8),
};

// AFFT_NO_ROW_SPLIT=1 hides the split-exchange rows (their lengths then take the multi-pass
// plans)
bool row_usable(const RowEntry &e)
{
    static const bool no_split = [] {
        const char *v = std::getenv("AFFT_NO_ROW_SPLIT");
        return v && v[0] == '1';
    }();
    return !(e.split && no_split);
}

const RowEntry *find_row(int dtype, int n, int variant)
{
    if (variant == kRowTransposed || variant == kRowResidual) {
        for (const RowEntry &e : kRows)
            if (e.dtype == dtype && e.n == n && (variant == kRowTransposed ? e.tr : e.res)) return &e;
        return nullptr;
    }
    const RowEntry *first = nullptr;
    int seen = 0;
    for (const RowEntry &e : kRows)
        if (e.dtype == dtype && e.n == n && !e.tr && !e.res && row_usable(e)) {
            if (!first) first = &e;
            if (seen++ == variant) return &e;
        }
    return variant == 0 ? first : nullptr;
}

} // namespace

int row_variants(int dtype, int n)
{
    int v = 0;
    for (const RowEntry &e : kRows)
        if (e.dtype == dtype && e.n == n && !e.tr && !e.res && row_usable(e)) ++v;
    return v;
}

// The default variant of a length for a batch of `bytes`: the three-waves-per-SIMD layout
// (variant 1) for Complex Float 4096, and for 1024 and 2048 from 128 MiB, whose tiles pair best
// with one tile per workgroup on large batches (16384 x 4096: 176.7 us against 178.9 us for
// variant 0; 128 MiB batches with four resident rounds 45.5-46.8 against 46.3-48.3 us).  It is also
// the fastest persistent layout of 4096 at 64 MiB (21.7 against 25.2 us), where 1024 and 2048 run
// variant 0 (24.2 / 25.5 against 25.0 / 25.8 us); variant 0 elsewhere.
int row_default_variant(int dtype, int n, int64_t bytes)
{
    if (dtype != 0 || row_variants(dtype, n) < 2) return 0;
    if (n == 4096) return 1;
    if ((n == 1024 || n == 2048) && bytes >= ((int64_t)128 << 20)) return 1;
    return 0;
}

// Whether a batch of these rows runs better with more workgroups than are resident: each one
// retires once its stores are issued and the dispatcher starts the next, whose loads overlap them.
// Measured on the compiled lengths up to 4096 Complex Float and 2048 Complex Double (the longer
// rows, one or two to a workgroup, lose: 8192 x 8192 236 -> 275 us, CD 4096 x 8192 189 -> 250 us).
bool row_grid_beyond_resident(int dtype, int n, int variant)
{
    const RowEntry *e = find_row(dtype, n, variant);
    if (!e || e->split || e->tr) return false;
    return dtype == 0 ? n <= 4096 : n <= 2048;
}

bool row_split(int dtype, int n)
{
    const RowEntry *e = find_row(dtype, n, 0);
    return e && e->split;
}

int row_block(int dtype, int n, int variant)
{
    const RowEntry *e = find_row(dtype, n, variant);
    return e ? e->block : 0;
}

size_t row_lds_bytes(int dtype, int n, int variant)
{
    const RowEntry *e = find_row(dtype, n, variant);
    if (!e) return 0;
    if (e->split) return (size_t)e->n * (dtype == 0 ? 4 : 8);
    const int C = e->rows;
    size_t elems = (size_t)C * lds_padded_len(e->n);
    if (e->tr) elems = std::max(elems, (size_t)e->n * (size_t)(C + 1)); // the regroup buffer
    return elems * (dtype == 0 ? 8 : 16);
}

int row_rows_per_tile(int dtype, int n, int variant)
{
    const RowEntry *e = find_row(dtype, n, variant);
    return e ? e->rows : 0;
}

int row_occupancy(int dtype, int n, int variant)
{
    const RowEntry *e = find_row(dtype, n, variant);
    if (!e) return 0;
    const size_t lds = row_lds_bytes(dtype, n, variant);
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(e->fn),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void *>(e->fn),
                                                     e->block, lds) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

hipError_t launch_rows(int dtype, int n, int variant, const RowArgs &a, int64_t grid, hipStream_t s)
{
    const RowEntry *e = find_row(dtype, n, variant);
    if (!e) return hipErrorInvalidValue;
    if (grid <= 0) return hipSuccess;
    const size_t lds = row_lds_bytes(dtype, n, variant);
    if (lds > 64 * 1024) {
        hipError_t err = hipFuncSetAttribute(reinterpret_cast<const void *>(e->fn),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(e->fn, dim3((unsigned)grid), dim3(e->block), lds, s, a);
    return hipGetLastError();
}

} // namespace afft
