// afft_internal.hpp -- launch descriptors shared by the planner (afft_plan.cpp) and the kernels,
// the launchers, and the device-side address helpers every batched kernel uses.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "afft_device.hpp"

namespace afft {

// Addressing of a batch of 1D transforms.  Transform j decomposes as
//   j0 = j % cnt0,  j1 = (j / cnt0) % cnt1,  j2 = j / (cnt0 * cnt1)
// and its element i lives at  j0*st0 + j1*st1 + j2*st2 + i*es  (in elements), or, with a
// gather table (ragged batches: rows of one length at arbitrary offsets), at
// j0*st0 + j1*st1 + toff[j2] + i*es.
struct Addr {
    int64_t cnt0, cnt1;
    int64_t st0, st1, st2;
    int64_t es;
    const int64_t *toff;
};

// offset of the first element of transform j (j = j0 + cnt0 * (j1 + cnt1 * j2))
__device__ __forceinline__ int64_t addr_base(const Addr &a, int64_t j)
{
    const int64_t j0 = j % a.cnt0;
    const int64_t q = j / a.cnt0;
    const int64_t j1 = q % a.cnt1;
    const int64_t j2 = q / a.cnt1;
    return j0 * a.st0 + j1 * a.st1 + (a.toff ? a.toff[j2] : j2 * a.st2);
}

__device__ __forceinline__ int64_t addr_level(const Addr &a, int64_t j, int level)
{
    return level == 0 ? j % a.cnt0 : (j / a.cnt0) % a.cnt1;
}

template <typename V>
__device__ __forceinline__ V power_lookup(const V *hi, const V *lo, int64_t e, int shift)
{
    const V a = hi[e >> shift];
    const V b = lo[e & ((int64_t(1) << shift) - 1)];
    return c_mul(a, b);
}

// Division of 0 <= x < 2^31 by a fixed d >= 1 as q = (x * m) >> sh (Granlund-Montgomery with
// s = ceil(log2 d), m = floor(2^(31+s) / d) + 1, sh = 31 + s: m*d lies in (2^(31+s), 2^(31+s) + d]
// and d <= 2^s, which makes the quotient exact for every 31-bit x; x*m < 2^63).
struct FastDiv {
    uint64_t m;
    int sh;
};

// A sign (-1)^(sum of the multi-index) on the elements of a row-major array of rank <= 3, applied
// on the load of a plan's first pass (centring the input) or on the store of its last pass: the
// fused form of DFT/Centre.hs.  div[a] divides by the stride of axis a (the product of the
// extents inside it); the array has fewer than 2^31 elements.
struct SignSpec {
    int on;
    int rank;
    int64_t ext[3];
    FastDiv div[3];
};

// parity of sum_a coord_a(off): with q_a = floor(off / stride_a), coord_a = q_a - ext_a * q_{a-1},
// so parity(coord_a) = parity(q_a) ^ (parity(ext_a) & parity(q_{a-1}))
__device__ __forceinline__ int sign_parity(const SignSpec &s, int64_t off)
{
    const uint64_t x = (uint64_t)off;
    int p = 0;
    uint64_t qprev = 0;
    for (int a = 0; a < s.rank; ++a) {
        const uint64_t q = (x * s.div[a].m) >> s.div[a].sh;
        p ^= (int)((q ^ ((uint64_t)(s.ext[a] & 1) & qprev)) & 1);
        qprev = q;
    }
    return p;
}

template <typename V> __device__ __forceinline__ V sign_apply(const SignSpec &s, int64_t off, V v)
{
    if (!s.on) return v;
    return sign_parity(s, off) ? c_make<V>(-v.x, -v.y) : v;
}

// One batched Stockham launch: ntrans transforms of length pt.n, C of them per workgroup.
struct FftArgs {
    const void *in;
    void *out;
    const void *tw;      // pass twiddles (element type), pass p at tw + pt.tw_off[p]
    const void *pw;      // W_n^m, m in [0, n), for the compile-time strided transforms, or null
    int64_t ntrans;
    Addr ain, aout;
    PassTable pt;
    int E;               // elements per thread (T = n / E threads per transform)
    int C;               // transforms per workgroup
    int lane_col;        // 1: the transform index varies fastest across lanes (strided axes)
    int lds_interleave;  // 1: LDS element (c, pos) at pad(pos)*C + c; 0: c*padded(n) + pad(pos)
    int conj_in;         // conjugate on load (direction +1)
    int conj_out;        // conjugate on store (direction +1)
    double scale;        // applied on store
    // four-step twiddle on store: out[k] *= W_L^{(jl * k) mod L}, jl = level-`post_level` index
    // of the transform (0 -> j0, 1 -> j1); W from the two-level table hi[e >> s] * lo[e & (2^s-1)]
    int64_t post_L;      // 0 = none
    int post_level;
    int post_shift;
    const void *post_hi;
    const void *post_lo;
    // the same twiddles as one precomputed grid, W_L^{jl k} at post_tab[k * post_ld + jl]: one
    // load per element, coalesced across the lanes of a strided tile (or null: the two-level table)
    const void *post_tab;
    int64_t post_ld;
    int xcd_remap;       // map workgroups to tiles XCD-contiguously (strided passes)
    int nt_in, nt_out;   // streaming loads / stores (plan input read once / final output)
    SignSpec sg_in, sg_out; // centring signs on load (first pass) / on store (last pass)
    int lanes;           // register kernel: lanes per transform (1; 2 or 4 exchange through DPP)
};

// Fused single-workgroup Bluestein (chirp-z) launch for a length n that is not 5-smooth:
// pt describes the padded length M (5-smooth, M >= 2n - 1).
struct BluesteinArgs {
    const void *in;
    void *out;
    const void *tw;      // pass twiddles for length M
    const void *pw;      // W_M^m, m in [0, M), for the compile-time transforms of M, or null
    const void *chirp;   // c[m] = exp(-i pi m^2 / n), m < n
    const void *filt;    // FFT_M(b) / M, b[m] = conj(c[|m|]) wrapped
    int64_t ntrans;
    int64_t n;
    Addr ain, aout;
    PassTable pt;
    int E, C;
    int lane_col, lds_interleave;
    int conj_in, conj_out;
    double scale;
    int xcd_remap;
    int nt_in, nt_out;
    SignSpec sg_in, sg_out;
};

// Pointwise stages of the multi-kernel Bluestein path (lengths whose padded size M exceeds one
// workgroup).  mode 0 (pre):  aux[j*M + m] = x_j[m] c[m] for m < n, 0 for n <= m < M
//                mode 1 (mid):  aux[j*M + m] = conj(aux[j*M + m] * filt[m])
//                mode 2 (post): y_j[m] = conj(aux[j*M + m]) c[m] for m < n
struct ChirpArgs {
    const void *in;
    void *out;
    const void *chirp;
    const void *filt;
    int64_t ntrans, n, M;
    Addr a;              // addressing of the caller's array (x for pre, y for post)
    int mode;
    int conj_in, conj_out;
    double scale;
    int nt_in, nt_out;
    SignSpec sg;         // pre: sign on the load of x; post: sign on the store of y
};
hipError_t launch_chirp(int dtype, const ChirpArgs &a, hipStream_t s);

// Persistent batch of contiguous rows of a length with a compiled layout (afft_rows.hip).
struct RowArgs {
    const void *in;
    void *out;
    const void *tw;      // W_N^m, m in [0, N), element type
    int64_t nrows;
    int conj_in;
    int conj_out;
    double scale;
    int nt_in, nt_out;
    const int64_t *roffs; // element offset of every row (ragged batches), or null: row * N
    SignSpec sg_in, sg_out;
    // transposed store (variant kRowTransposed): row r = (o, k1) of n1 rows per transform, its
    // element k2 stored at (tr_toff ? tr_toff[o] : o * N * n1) + k1 + n1 * k2
    int64_t tr_n1;
    const int64_t *tr_toff;
    // rotated tiles (tr_rot, needs C | n1): tile slot u of transform o holds row
    // k1 = (u - s_o) mod n1, s_o = the element address of X_o[0] modulo C, so that each tile's
    // stores fill aligned runs of C elements wherever the transform starts
    int tr_rot;
    // residual loads (variant kRowResidual): the rows are alpha (in - in2), in2 laid out as in
    const void *in2;
    double alpha;
};

// afft_rows.hip: number of compiled variants for (dtype, n) (0 = none), their launch shape,
// LDS bytes, rows per tile, resident workgroups per CU, and the launcher.
constexpr int kRowTransposed = -2; // the variant index of a length's transposed-store form
constexpr int kRowResidual = -3;   // the variant index of a length's residual-load form
int row_variants(int dtype, int n);
// the variant a plan takes for a batch of `bytes` when neither AFFT_ROW_VARIANT nor the autotuner
// chooses one
int row_default_variant(int dtype, int n, int64_t bytes);
// true when a large batch of these rows runs faster with more workgroups than are resident
bool row_grid_beyond_resident(int dtype, int n, int variant);
// true when the default layout of (dtype, n) is a split-exchange one (one row per workgroup)
bool row_split(int dtype, int n);
int row_block(int dtype, int n, int variant);
size_t row_lds_bytes(int dtype, int n, int variant);
int row_rows_per_tile(int dtype, int n, int variant);
int row_occupancy(int dtype, int n, int variant);
hipError_t launch_rows(int dtype, int n, int variant, const RowArgs &a, int64_t grid, hipStream_t s);

// Batched strided transpose (afft_cols.hip): matrix o = o1 + nb1 * o2 of nbatch, element (r, c)
// at in + o1 in_bs1 + o2 in_bs2 + r in_rs + c, stored transposed with (c, r) at
// out + o1 out_bs1 + (toff ? toff[o2] : o2 out_bs2) + c out_rs + r, conjugated and scaled.
struct TransposeArgs {
    const void *in;
    void *out;
    int64_t rows, cols, nbatch;
    int64_t nb1, in_bs1, in_bs2, in_rs, out_bs1, out_bs2, out_rs;
    const int64_t *toff;
    int conj_out;
    double scale;
    int nt_out;
    int xcd_remap;
    SignSpec sg_out;
};
hipError_t launch_transpose(int dtype, const TransposeArgs &a, hipStream_t s);

// Launchers (afft_kernels.hip).  They return a hipError_t.
// register-resident strided transforms (afft_cols.hip): one thread per transform of length
// a.E, kRegColBlock threads per workgroup, no LDS
bool regcol_supported(int dtype, int64_t n);
int regcol_block();
hipError_t launch_regcol(int dtype, const FftArgs &a, int64_t grid, hipStream_t s);
hipError_t launch_stockham(int dtype, const FftArgs &a, int block, int64_t grid, size_t lds,
                           hipStream_t s);
hipError_t launch_bluestein(int dtype, const BluesteinArgs &a, int block, int64_t grid, size_t lds,
                            hipStream_t s);
// FP64 twiddle generation, rounded to the element type.
hipError_t gen_pass_twiddles(int dtype, void *out, const PassTable &pt, hipStream_t s);
// out[i] = W_L^{i * step} for i < count (two-level four-step tables).
// out[q * cols + j] = W_L^{q j}, q < rows, j < cols, rows * cols <= L (FP64, rounded)
hipError_t gen_twiddle_grid(int dtype, void *out, int64_t L, int64_t rows, int64_t cols, hipStream_t s);
hipError_t gen_power_table(int dtype, void *out, int64_t L, int64_t step, int64_t count,
                           hipStream_t s);
// out[m] = exp(-i pi (m^2 mod 2n) / n) for m < n (chirp), in the element type.
hipError_t gen_chirp(int dtype, void *out, int64_t n, hipStream_t s);
// out (complex double, length M) = b[m]: conj(chirp[m]) for m < n, conj(chirp[M-m]) for
// M - m < n, 0 otherwise, computed in FP64.
hipError_t gen_bluestein_kernel_f64(void *out, int64_t n, int64_t M, hipStream_t s);
// out (element type) = in (complex double) * scale, count elements.
hipError_t convert_scale_from_f64(int dtype, void *out, const void *in, int64_t count, double scale,
                                  hipStream_t s);
// y = x * h (elementwise, complex), count elements; h indexed modulo hcount.
hipError_t pointwise_mul(int dtype, void *y, const void *x, const void *h, int64_t count,
                         int64_t hcount, int conj_h, double scale, hipStream_t s);

// The elements-per-thread values that have kernel instantiations.
extern const int kElemsPerThread[];
extern const int kNumElemsPerThread;

} // namespace afft
