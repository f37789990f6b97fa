// afft_device.hpp -- device-side building blocks of the gfx950 FFT kernels.
//
// Complex values are HIP float2/double2 (re, im): the interleaved Vec2 layout of
// Complex Float / Complex Double (src/Data/Array/Accelerate/Math/FFT/Type.hs:32).
//
// Every kernel computes the *forward* DFT X[k] = sum_n x[n] exp(-2 pi i nk/N).  The opposite
// direction (Reverse / Inverse, FFT/Mode.hs:21-26) is conj(F(conj x)): the conjugations are
// folded into the load and the store, so one set of butterflies serves both directions.
//
// Stockham autosort formulation (Govindaraju et al., "High performance discrete Fourier
// transforms on graphics processors", SC'08): pass p of radix R with Ns = product of the earlier
// radices reads butterfly b's inputs at b + r*N/R, multiplies input r by W_{Ns R}^{r (b mod Ns)},
// applies DFT_R and writes output q to (b - b mod Ns)*R + (b mod Ns) + q*Ns.
//
// Register layout.  A transform of length N is owned by T = N/E threads, each holding E
// elements.  Thread t holds element t + i*T in reg[i] ("natural" layout).  For a pass of radix R
// (R divides E), butterfly j of thread t is b = t + j*T and uses reg[j + r*E/R]; the last pass
// writes exactly the natural layout, so the first pass loads straight from global memory and the
// last pass stores straight to it, and only the passes in between exchange through LDS.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <utility>

namespace afft {

template <typename Real> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };

template <typename V> __device__ __forceinline__ V c_make(decltype(V::x) re, decltype(V::x) im)
{
    V v;
    v.x = re;
    v.y = im;
    return v;
}
template <typename V> __device__ __forceinline__ V c_add(V a, V b) { return c_make<V>(a.x + b.x, a.y + b.y); }
template <typename V> __device__ __forceinline__ V c_sub(V a, V b) { return c_make<V>(a.x - b.x, a.y - b.y); }
template <typename V> __device__ __forceinline__ V c_mul(V a, V b)
{
    return c_make<V>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename V> __device__ __forceinline__ V c_mulc(V a, V b)
{
    return c_make<V>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename V> __device__ __forceinline__ V c_conj(V a) { return c_make<V>(a.x, -a.y); }
// a * (-i)
template <typename V> __device__ __forceinline__ V c_mul_negi(V a) { return c_make<V>(a.y, -a.x); }
template <typename V> __device__ __forceinline__ V c_scale(V a, decltype(V::x) s) { return c_make<V>(a.x * s, a.y * s); }

// Streaming (non-temporal) accesses for data read or written exactly once by a plan: they do
// not displace the intermediate arrays a multi-pass plan keeps in L2 / Infinity Cache.
template <typename V> __device__ __forceinline__ V c_load_nt(const V *p)
{
    typedef decltype(V::x) Real;
    typedef Real vec2 __attribute__((ext_vector_type(2)));
    const vec2 v = __builtin_nontemporal_load(reinterpret_cast<const vec2 *>(p));
    return c_make<V>(v.x, v.y);
}
template <typename V> __device__ __forceinline__ void c_store_nt(V *p, V a)
{
    typedef decltype(V::x) Real;
    typedef Real vec2 __attribute__((ext_vector_type(2)));
    vec2 v;
    v.x = a.x;
    v.y = a.y;
    __builtin_nontemporal_store(v, reinterpret_cast<vec2 *>(p));
}
template <typename V> __device__ __forceinline__ V c_load(const V *p, bool nt) { return nt ? c_load_nt(p) : *p; }

// Pointers known to address global memory: loads and stores through them compile to global_*
// instructions, not the flat ones a generic pointer gives inside a non-inlined function (flat
// accesses also count against the LDS wait counter, so every LDS wait and barrier after them
// would wait for them too).  The element moves as a builtin two-lane vector.
template <typename V> using GIn = const V __attribute__((address_space(1)));
template <typename V> using GOut = V __attribute__((address_space(1)));

template <typename V> __device__ __forceinline__ V ld_elem(const V *p) { return *p; }

template <typename V> __device__ __forceinline__ V ld_elem(GIn<V> *p)
{
    typedef decltype(V::x) R2 __attribute__((ext_vector_type(2)));
    const R2 r = *reinterpret_cast<GIn<R2> *>(p);
    return c_make<V>(r.x, r.y);
}

template <typename V> __device__ __forceinline__ void st_elem(GOut<V> *p, V v)
{
    typedef decltype(V::x) R2 __attribute__((ext_vector_type(2)));
    R2 r;
    r.x = v.x;
    r.y = v.y;
    *reinterpret_cast<GOut<R2> *>(p) = r;
}
template <typename V> __device__ __forceinline__ void c_store(V *p, V a, bool nt)
{
    if (nt) c_store_nt(p, a);
    else *p = a;
}

// Workgroup -> tile order that keeps consecutive tiles on one XCD.  The dispatcher hands
// workgroup b to XCD b mod 8; tiles that share cache lines (adjacent column tiles of a strided
// pass) must meet in the same L2, so XCD x takes the contiguous tile range [x G/8, (x+1) G/8).
__device__ __forceinline__ int64_t xcd_tile(int64_t b, int64_t G)
{
    constexpr int kXcd = 8;
    const int64_t full = (G / kXcd) * kXcd;
    if (b >= full) return b;
    return (b % kXcd) * (G / kXcd) + b / kXcd;
}

// ----------------------------------------------------------------------------------------------
// Forward DFT_R butterflies on v[0..R-1], in place.

template <int R> struct dft;

template <> struct dft<1> {
    template <typename V> __device__ __forceinline__ static void run(V *) {}
};

template <> struct dft<2> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        V a = v[0], b = v[1];
        v[0] = c_add(a, b);
        v[1] = c_sub(a, b);
    }
};

template <> struct dft<3> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        using Real = decltype(V::x);
        const Real h = (Real)0.86602540378443864676372317075293618; // sqrt(3)/2
        V t1 = c_add(v[1], v[2]);
        V d = c_sub(v[1], v[2]);
        V t2 = c_make<V>(v[0].x - (Real)0.5 * t1.x, v[0].y - (Real)0.5 * t1.y);
        V m = c_make<V>(h * d.y, -h * d.x); // -i * (sqrt3/2) * d
        v[0] = c_add(v[0], t1);
        v[1] = c_add(t2, m);
        v[2] = c_sub(t2, m);
    }
};

template <> struct dft<4> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        V s0 = c_add(v[0], v[2]), d0 = c_sub(v[0], v[2]);
        V s1 = c_add(v[1], v[3]), d1 = c_mul_negi(c_sub(v[1], v[3]));
        v[0] = c_add(s0, s1);
        v[2] = c_sub(s0, s1);
        v[1] = c_add(d0, d1);
        v[3] = c_sub(d0, d1);
    }
};

template <> struct dft<5> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        using Real = decltype(V::x);
        const Real c1 = (Real)0.30901699437494742410229341718281906;  // cos(2pi/5)
        const Real c2 = (Real)-0.80901699437494742410229341718281906; // cos(4pi/5)
        const Real s1 = (Real)0.95105651629515357211643933337938214;  // sin(2pi/5)
        const Real s2 = (Real)0.58778525229247312916870595463907277;  // sin(4pi/5)
        V t1 = c_add(v[1], v[4]), d1 = c_sub(v[1], v[4]);
        V t2 = c_add(v[2], v[3]), d2 = c_sub(v[2], v[3]);
        V a1 = c_make<V>(v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y);
        V a2 = c_make<V>(v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y);
        V b1 = c_make<V>(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y);
        V b2 = c_make<V>(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y);
        v[0] = c_make<V>(v[0].x + t1.x + t2.x, v[0].y + t1.y + t2.y);
        // y1 = a1 - i b1, y4 = a1 + i b1, y2 = a2 - i b2, y3 = a2 + i b2
        v[1] = c_make<V>(a1.x + b1.y, a1.y - b1.x);
        v[4] = c_make<V>(a1.x - b1.y, a1.y + b1.x);
        v[2] = c_make<V>(a2.x + b2.y, a2.y - b2.x);
        v[3] = c_make<V>(a2.x - b2.y, a2.y + b2.x);
    }
};

template <> struct dft<8> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        using Real = decltype(V::x);
        const Real r = (Real)0.70710678118654752440084436210484904; // 1/sqrt(2)
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        dft<4>::run(e);
        dft<4>::run(o);
        // o[k] *= W8^k: W8 = (1 - i)/sqrt2, W8^2 = -i, W8^3 = (-1 - i)/sqrt2
        V o1 = c_make<V>(r * (o[1].x + o[1].y), r * (o[1].y - o[1].x));
        V o2 = c_mul_negi(o[2]);
        V o3 = c_make<V>(r * (o[3].y - o[3].x), -r * (o[3].x + o[3].y));
        v[0] = c_add(e[0], o[0]);
        v[4] = c_sub(e[0], o[0]);
        v[1] = c_add(e[1], o1);
        v[5] = c_sub(e[1], o1);
        v[2] = c_add(e[2], o2);
        v[6] = c_sub(e[2], o2);
        v[3] = c_add(e[3], o3);
        v[7] = c_sub(e[3], o3);
    }
};

template <> struct dft<16> {
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        using Real = decltype(V::x);
        const Real c8 = (Real)0.92387953251128675612818318939678829;  // cos(pi/8)
        const Real s8 = (Real)0.38268343236508977172845998403039887;  // sin(pi/8)
        const Real r2 = (Real)0.70710678118654752440084436210484904;  // 1/sqrt(2)
        // n = r1 + 4 n2, k = k1 + 4 k2
        V a[4][4];
#pragma unroll
        for (int r1 = 0; r1 < 4; ++r1) {
            V t[4] = {v[r1], v[r1 + 4], v[r1 + 8], v[r1 + 12]};
            dft<4>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < 4; ++k1) a[r1][k1] = t[k1];
        }
        // twiddles W16^(r1*k1)
        const V w1 = c_make<V>(c8, -s8), w2 = c_make<V>(r2, -r2), w3 = c_make<V>(s8, -c8);
        const V w6 = c_make<V>(-r2, -r2), w9 = c_make<V>(-c8, s8);
        a[1][1] = c_mul(a[1][1], w1);
        a[1][2] = c_mul(a[1][2], w2);
        a[1][3] = c_mul(a[1][3], w3);
        a[2][1] = c_mul(a[2][1], w2);
        a[2][2] = c_mul_negi(a[2][2]);
        a[2][3] = c_mul(a[2][3], w6);
        a[3][1] = c_mul(a[3][1], w3);
        a[3][2] = c_mul(a[3][2], w6);
        a[3][3] = c_mul(a[3][3], w9);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            V t[4] = {a[0][k1], a[1][k1], a[2][k1], a[3][k1]};
            dft<4>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = t[k2];
        }
    }
};

// ----------------------------------------------------------------------------------------------
// Composite radices R = R1 * R2 in registers (used for 32 and 64).
//   n = R2 n1 + n2, k = k1 + R1 k2:  X[k] = sum_n2 W_R^{n2 k1} W_R2^{n2 k2} sum_n1 x[n] W_R1^{n1 k1}
// The twiddles W_R^m are compile-time constants: cos and sin of 2 pi m / R evaluated by a
// constexpr Taylor series after reduction to [0, pi/4] (correct to double precision).

constexpr double kPi = 3.14159265358979323846264338327950288;

__host__ __device__ constexpr double taylor_sin(double x)
{
    double term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
__host__ __device__ constexpr double taylor_cos(double x)
{
    double term = 1, sum = 1;
    for (int k = 1; k < 14; ++k) {
        term *= -x * x / ((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
// quarter-period pieces: x in [0, pi/2], reduced to [0, pi/4] by cos x = sin(pi/2 - x)
__host__ __device__ constexpr double cos_q(double x) { return x <= kPi / 4 ? taylor_cos(x) : taylor_sin(kPi / 2 - x); }
__host__ __device__ constexpr double sin_q(double x) { return x <= kPi / 4 ? taylor_sin(x) : taylor_cos(kPi / 2 - x); }
// cos and sin of 2 pi m / R, m reduced exactly in integers first
__host__ __device__ constexpr double cos2pi(int m, int R)
{
    m %= R;
    if (m < 0) m += R;
    if (2 * m > R) m = R - m;                   // cos(2 pi - x) = cos x
    const double th = 2 * kPi * m / R;          // [0, pi]
    return th <= kPi / 2 ? cos_q(th) : -cos_q(kPi - th);
}
__host__ __device__ constexpr double sin2pi(int m, int R)
{
    m %= R;
    if (m < 0) m += R;
    double sg = 1;
    if (2 * m > R) {                            // sin(2 pi - x) = -sin x
        m = R - m;
        sg = -1;
    }
    const double th = 2 * kPi * m / R;
    return sg * (th <= kPi / 2 ? sin_q(th) : sin_q(kPi - th));
}

// v *= W_R^m (forward root exp(-2 pi i m / R)), trivial factors without a multiply
template <int R, int M, typename V> __device__ __forceinline__ V twiddle_const(V v)
{
    using Real = decltype(V::x);
    constexpr int m = ((M % R) + R) % R;
    if constexpr (m == 0) return v;
    else if constexpr (4 * m == R) return c_mul_negi(v);
    else if constexpr (2 * m == R) return c_make<V>(-v.x, -v.y);
    else if constexpr (4 * m == 3 * R) return c_make<V>(-v.y, v.x);
    else {
        constexpr Real c = (Real)cos2pi(m, R);
        constexpr Real sn = (Real)-sin2pi(m, R);
        return c_make<V>(v.x * c - v.y * sn, v.x * sn + v.y * c);
    }
}

template <int R> struct dft;

template <int R1, int R2> struct dft_split {
    static constexpr int R = R1 * R2;
    template <int N2, int K1, typename V> __device__ __forceinline__ static void tw(V (&a)[R2][R1])
    {
        a[N2][K1] = twiddle_const<R, N2 * K1>(a[N2][K1]);
    }
    template <typename V> __device__ __forceinline__ static void run(V *v)
    {
        V a[R2][R1];
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) {
            V t[R1];
#pragma unroll
            for (int n1 = 0; n1 < R1; ++n1) t[n1] = v[R2 * n1 + n2];
            dft<R1>::run(t);
#pragma unroll
            for (int k1 = 0; k1 < R1; ++k1) a[n2][k1] = t[k1];
        }
        twiddle_all(a, std::make_integer_sequence<int, R1 * R2>{});
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) {
            V t[R2];
#pragma unroll
            for (int n2 = 0; n2 < R2; ++n2) t[n2] = a[n2][k1];
            dft<R2>::run(t);
#pragma unroll
            for (int k2 = 0; k2 < R2; ++k2) v[k1 + R1 * k2] = t[k2];
        }
    }
    template <typename V, int... Is>
    __device__ __forceinline__ static void twiddle_all(V (&a)[R2][R1], std::integer_sequence<int, Is...>)
    {
        (tw<Is / R1, Is % R1>(a), ...);
    }
};

template <> struct dft<32> {
    template <typename V> __device__ __forceinline__ static void run(V *v) { dft_split<8, 4>::run(v); }
};
template <> struct dft<64> {
    template <typename V> __device__ __forceinline__ static void run(V *v) { dft_split<8, 8>::run(v); }
};

// ----------------------------------------------------------------------------------------------
// Pass descriptors (shared by host and device).

constexpr int kMaxPasses = 24;

// Largest workgroup a kernel holding E complex elements per thread is compiled for.  The launch
// bound sets the VGPR budget (512 per SIMD lane slot / waves per SIMD): 1024 threads -> 128
// VGPRs, 512 -> 256, 256 -> 512.  One FP32 complex is 2 VGPRs, one FP64 complex 4; the
// butterflies need about twice the data registers.
__host__ __device__ constexpr int max_block_threads(bool f64, int E)
{
    return (f64 ? 2 * E : E) <= 16 ? 1024 : ((f64 ? 2 * E : E) <= 32 ? 512 : 256);
}

struct PassTable {
    int n;                       // transform length
    int npass;                   // number of Stockham passes
    int radix[kMaxPasses];       // radix of each pass
    int tw_off[kMaxPasses];      // offset (in elements) of the pass's twiddles, stride R per k
};

// LDS address of element `pos` of the transform whose LDS base is `base`: one padding element
// after every 16 spreads the strided Stockham writes over the banks.
__device__ __forceinline__ int lds_pad(int pos) { return pos + (pos >> 4); }
__host__ __device__ constexpr int lds_padded_len(int n) { return n + (n >> 4) + 1; }

// Twiddle-multiply + DFT_R for every butterfly the thread owns in the current pass.
template <int E, int R, typename V>
__device__ __forceinline__ void pass_butterflies(V (&reg)[E], const V *__restrict__ tw, int Ns,
                                                 int t, int T)
{
    constexpr int B = E / R;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        V v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) v[r] = reg[j + r * B];
        if (Ns > 1) {
            const int b = t + j * T;
            const int k = b % Ns;
            const V *w = tw + (int64_t)k * R;
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = c_mul(v[r], w[r]);
        }
        dft<R>::run(v);
#pragma unroll
        for (int r = 0; r < R; ++r) reg[j + r * B] = v[r];
    }
}

// Scatter the pass outputs to their Stockham positions in LDS.
// lds(pos) = lds_base[lds_pad(pos) * es]  (es = 1 for a private row, C for interleaved columns)
template <int E, int R, typename V>
__device__ __forceinline__ void pass_scatter(const V (&reg)[E], V *lds_base, int es, int Ns, int t,
                                             int T)
{
    constexpr int B = E / R;
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const int b = t + j * T;
        const int k = b % Ns;
        const int dst = (b - k) * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds_base[lds_pad(dst + r * Ns) * es] = reg[j + r * B];
    }
}

template <int E, typename V>
__device__ __forceinline__ void natural_read(V (&reg)[E], const V *lds_base, int es, int t, int T)
{
#pragma unroll
    for (int i = 0; i < E; ++i) reg[i] = lds_base[lds_pad(t + i * T) * es];
}

template <int E, typename V>
__device__ __forceinline__ void natural_write(const V (&reg)[E], V *lds_base, int es, int t, int T)
{
#pragma unroll
    for (int i = 0; i < E; ++i) lds_base[lds_pad(t + i * T) * es] = reg[i];
}

// One pass of radix R, dispatched at run time.  Only radices that divide E are instantiated.
template <int E, typename V>
__device__ __forceinline__ void pass_dispatch(int R, V (&reg)[E], const V *__restrict__ tw, int Ns,
                                              int t, int T)
{
    switch (R) {
    case 2: if constexpr (E % 2 == 0) pass_butterflies<E, 2>(reg, tw, Ns, t, T); break;
    case 3: if constexpr (E % 3 == 0) pass_butterflies<E, 3>(reg, tw, Ns, t, T); break;
    case 4: if constexpr (E % 4 == 0) pass_butterflies<E, 4>(reg, tw, Ns, t, T); break;
    case 5: if constexpr (E % 5 == 0) pass_butterflies<E, 5>(reg, tw, Ns, t, T); break;
    case 8: if constexpr (E % 8 == 0) pass_butterflies<E, 8>(reg, tw, Ns, t, T); break;
    case 16: if constexpr (E % 16 == 0) pass_butterflies<E, 16>(reg, tw, Ns, t, T); break;
    default: break;
    }
}

template <int E, typename V>
__device__ __forceinline__ void scatter_dispatch(int R, const V (&reg)[E], V *lds_base, int es,
                                                 int Ns, int t, int T)
{
    switch (R) {
    case 2: if constexpr (E % 2 == 0) pass_scatter<E, 2>(reg, lds_base, es, Ns, t, T); break;
    case 3: if constexpr (E % 3 == 0) pass_scatter<E, 3>(reg, lds_base, es, Ns, t, T); break;
    case 4: if constexpr (E % 4 == 0) pass_scatter<E, 4>(reg, lds_base, es, Ns, t, T); break;
    case 5: if constexpr (E % 5 == 0) pass_scatter<E, 5>(reg, lds_base, es, Ns, t, T); break;
    case 8: if constexpr (E % 8 == 0) pass_scatter<E, 8>(reg, lds_base, es, Ns, t, T); break;
    case 16: if constexpr (E % 16 == 0) pass_scatter<E, 16>(reg, lds_base, es, Ns, t, T); break;
    default: break;
    }
}

// Ordering of an LDS exchange private to one wavefront (all threads of the transform in one
// wavefront): a wavefront's LDS instructions execute in order, so wavefront-scope ordering is
// enough and no workgroup barrier is needed.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

// The whole forward transform of the rows held in registers (natural layout in, natural layout
// out).  All threads of the workgroup must call it (it contains barriers); threads that own no
// transform (tail of the batch) still take part.  `lds_base`/`es` address this thread's
// transform in LDS; `tw` holds the pass twiddles (pass p at tw + pt.tw_off[p]).
// `on` = false for threads of the workgroup that own no transform slot: they take part in the
// barriers but never touch LDS.
// `wave` (uniform) = every transform's threads lie in one wavefront and its LDS slot is private
// to it: the exchanges then synchronise at wavefront scope instead of with workgroup barriers.
template <int E, typename V>
__device__ __forceinline__ void fft_core(V (&reg)[E], V *lds_base, int es, const PassTable &pt,
                                         const V *__restrict__ tw, int t, int T, bool on = true,
                                         bool wave = false)
{
    int Ns = 1;
    for (int p = 0; p < pt.npass; ++p) {
        const int R = pt.radix[p];
        pass_dispatch<E>(R, reg, tw + pt.tw_off[p], Ns, t, T);
        if (p + 1 < pt.npass) {
            if (wave) wave_sync(); else __syncthreads(); // the previous exchange has been read
            if (on) scatter_dispatch<E>(R, reg, lds_base, es, Ns, t, T);
            if (wave) wave_sync(); else __syncthreads();
            if (on) natural_read<E>(reg, lds_base, es, t, T);
        }
        Ns *= R;
    }
}

// ----------------------------------------------------------------------------------------------
// Compile-time Stockham helpers (the row kernels of afft_rows.hip and the column kernels of
// afft_spectral.hip): workgroup LDS barrier, radix lists, a compile-time loop, and the twiddle
// powers of one butterfly formed from two loaded bases.

__device__ __forceinline__ void lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int... Rs> struct Radices {
    static constexpr int NP = sizeof...(Rs);
    static constexpr int R[NP] = {Rs...};
    static constexpr int ns(int p)
    {
        int v = 1;
        for (int i = 0; i < p; ++i) v *= R[i];
        return v;
    }
};

template <int P> struct IC {
    static constexpr int value = P;
};

template <typename F, int... Ps>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, Ps...>)
{
    (f(IC<Ps>{}), ...);
}
template <int N, typename F> __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// w[r] = W^{r k} for r = 1..R-1 from b1 = W^k and b4 = W^{4k} (b4 used when R > 4).
template <int R, typename V> __device__ __forceinline__ void derive_twiddles(V *w, V b1, V b4)
{
    w[1] = b1;
    if constexpr (R > 2) w[2] = c_mul(b1, b1);
    if constexpr (R > 3) w[3] = c_mul(w[2], b1);
    if constexpr (R == 5) w[4] = c_mul(w[2], w[2]);
    if constexpr (R > 5) {
        w[4] = b4;
#pragma unroll
        for (int r = 5; r < R; ++r) {
            const int q = r & ~3, m = r & 3;
            if (m == 0) w[r] = c_mul(w[q - 4], b4);
            else w[r] = c_mul(w[q], w[m]);
        }
    }
}

// v[r] *= W^{r k} for r = 1..R-1, the powers formed on the fly from b1 = W^k, b4 = W^{4k}
// (the same products as derive_twiddles, with only the bases w1..w3, w4, w8, w12 kept live).
template <int R, typename V> __device__ __forceinline__ void apply_twiddles(V *v, V b1, V b4)
{
    v[1] = c_mul(v[1], b1);
    if constexpr (R > 2) {
        const V w2 = c_mul(b1, b1);
        v[2] = c_mul(v[2], w2);
        if constexpr (R > 3) {
            const V w3 = c_mul(w2, b1);
            v[3] = c_mul(v[3], w3);
            if constexpr (R == 5) v[4] = c_mul(v[4], c_mul(w2, w2));
            if constexpr (R > 5) {
                V w4 = b4;
#pragma unroll
                for (int q = 4; q < R; q += 4) {
                    if (q > 4) w4 = c_mul(w4, b4); // W^{q k}
                    v[q] = c_mul(v[q], w4);
                    v[q + 1] = c_mul(v[q + 1], c_mul(w4, b1));
                    v[q + 2] = c_mul(v[q + 2], c_mul(w4, w2));
                    v[q + 3] = c_mul(v[q + 3], c_mul(w4, w3));
                }
            }
        }
    }
}

// A transform of length N = R0 R1 ... with E elements per thread in natural layout (reg[i] =
// element t + i T, T = N / E) whose exchanges go through LDS at base[e * es]: the Stockham passes
// of afft_rows, with every radix, stride and butterfly index a compile-time constant and the
// twiddle bases (W^k and W^4k of each pass, from the table W_N^m) loaded once by load().  The
// exchanges synchronise the whole workgroup, so the T threads of a transform may sit in any
// waves (column tiles interleave the transforms across lanes).
template <typename V, int N, int E, int... Rs> struct CtStockham {
    using RL = Radices<Rs...>;
    template <bool PAD> static __device__ __forceinline__ int pos(int p) { return PAD ? lds_pad(p) : p; }
    static constexpr int T = N / E;
    static constexpr int NP = RL::NP;
    static_assert(T * E == N, "E must divide N");
    V tb[NP > 0 ? NP : 1][2 * E];

    __device__ __forceinline__ void load(const V *__restrict__ tw, int t) { load_from(tw, t); }
    __device__ __forceinline__ void load(GIn<V> *tw, int t) { load_from(tw, t); }

    template <typename P> __device__ __forceinline__ void load_from(P tw, int t)
    {
        static_for<NP>([&](auto pc) {
            constexpr int p = decltype(pc)::value;
            constexpr int R = RL::R[p];
            constexpr int Ns = RL::ns(p);
            if constexpr (Ns > 1) {
                constexpr int B = E / R;
                constexpr int NB = (T % Ns == 0 || Ns * R == N) ? 1 : B;
                constexpr int STEP = N / (Ns * R);
#pragma unroll
                for (int jb = 0; jb < NB; ++jb) {
                    const int k = (t + jb * T) % Ns;
                    tb[p][2 * jb] = ld_elem(tw + k * STEP);
                    tb[p][2 * jb + 1] = (R > 5) ? ld_elem(tw + 4 * k * STEP) : tb[p][2 * jb];
                }
            }
        });
    }

    // PAD: positions pass through lds_pad (a private contiguous slot, es = 1); interleaved tiles
    // (es = C) measured the same without it.  WAVE: the T threads of the transform sit in one
    // wavefront with a private slot, so the exchanges order at wavefront scope (no barrier).
    // FRESH: each exchange recomputes its LDS positions from the thread index (passed through an
    // empty asm), instead of the compiler computing every pass's positions up front and keeping
    // them live across the transform.
    template <bool PAD = false, bool WAVE = false, bool FRESH = false>
    __device__ __forceinline__ void run(V (&reg)[E], V *base, int es, int t0) const
    {
        static_assert(!WAVE || (T <= 64 && 64 % T == 0), "wavefront-private exchanges");
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
            if constexpr (p + 1 < NP) {
                int t = t0;
                if constexpr (FRESH) __asm__ volatile("" : "+v"(t));
                // the previous exchange has been read by every thread
                if constexpr (WAVE) wave_sync();
                else lds_sync();
#pragma unroll
                for (int j = 0; j < B; ++j) {
                    const int b = t + j * T;
                    const int k = b % Ns;
                    const int dst = (b - k) * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) base[pos<PAD>(dst + r * Ns) * es] = reg[j + r * B];
                }
                if constexpr (WAVE) wave_sync();
                else lds_sync();
#pragma unroll
                for (int i = 0; i < E; ++i) reg[i] = base[pos<PAD>(t + i * T) * es];
            }
        });
    }
};

} // namespace afft
