/*
 * afft_oracle.c -- CPU restatement of the Data.Array.Accelerate.Math.FFT hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  The shipped library never links or calls this file.  Only the
 * test-suite, the smoke check and the CPU-baseline leg of the benchmark load it, and only as
 * the checker (or as the CPU path timed beside the GPU), never as the thing measured or shipped.
 *
 * What it restates (citations are relative to the accelerate-fft v1.3.0.0 source tree):
 *
 *   Direction and sign      src/Data/Array/Accelerate/Math/FFT/Mode.hs:15-26
 *                           Forward = sign -1, Reverse = +1 unnormalised, Inverse = +1 then / N.
 *   Plugin contract         src/Data/Array/Accelerate/Math/FFT/LLVM/Native.hs:49-102 and
 *                           LLVM/Native/Base.hs:42-47: FFTW guru DFT, out-of-place, input
 *                           preserved, unnormalised, sign DFTForward/DFTBackward, over axes
 *                           [rank-1] (fft) or all axes (fft1D/fft2D/fft3D).
 *   Inverse scale           src/Data/Array/Accelerate/Math/FFT.hs:69,83 (innermost extent for
 *                           fft), :98,110 (length for fft1D), :125,141,155,172 (total size for
 *                           fft2D/fft3D).  The front-end divides after the plugin returns.
 *   Definition (brute)      src/Data/Array/Accelerate/Math/DFT.hs:68-111 (sum over n of
 *                           x[n] * root[(k*n) mod l]) with the roots of DFT/Roots.hs:26-51
 *                           (cos t - i sin t forward, cos t + i sin t inverse, t = 2 pi i / n).
 *
 * The FFTW library the reference binds (through the Hackage `fft` >= 0.1.8 package,
 * accelerate-fft.cabal:75) is absent from this image; its published algorithm is the DFT
 * definition above, which is what this file computes:
 *
 *   orc_dft_brute   O(N^2) definition in long double with the exact integer reduction
 *                   (j*k) mod N, the independent check of everything else here.
 *   orc_fft_*       recursive mixed-radix decimation-in-time Cooley-Tukey in double (or float
 *                   for the CPU-baseline timing), twiddles computed in long double and rounded,
 *                   any prime factor handled by a direct O(p^2) combine accumulated in long
 *                   double.  Deliberately a different algorithm from the GPU kernels (which are
 *                   iterative Stockham) so the two cannot share a mistake.
 *   orc_fft_nd      applies the 1D transform along a chosen set of axes (row-major,
 *                   innermost fastest, Accelerate's Z:.z:.y:.x order).
 *   orc_gen_*       the counter-based SplitMix64 input generator used for seeded inputs.
 *
 * Data layout: interleaved (re, im) pairs, the Vec2 layout of Complex e
 * (src/Data/Array/Accelerate/Math/FFT/Type.hs:32).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2

/* ------------------------------------------------------------------------------------------ */
/* Counter-based SplitMix64 generator: value i of stream `seed` depends only on (seed, i).     */

static inline uint64_t orc_mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static inline uint64_t orc_counter(uint64_t seed, uint64_t i)
{
    return orc_mix64(orc_mix64(seed) + i);
}

/* out[i] = uniform in [-1, 1): 53 random bits mapped to [0,1), then 2u - 1 (exact in double). */
void orc_gen_uniform(uint64_t seed, int64_t count, double *out)
{
    for (int64_t i = 0; i < count; ++i) {
        uint64_t h = orc_counter(seed, (uint64_t)i);
        double u = (double)(h >> 11) * 0x1.0p-53;
        out[i] = 2.0 * u - 1.0;
    }
}

/* out[2j], out[2j+1] = a standard-normal pair by Box-Muller from counters 2j and 2j+1. */
void orc_gen_normal(uint64_t seed, int64_t count, double *out)
{
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t j = 0; 2 * j < count; ++j) {
        uint64_t h1 = orc_counter(seed, (uint64_t)(2 * j));
        uint64_t h2 = orc_counter(seed, (uint64_t)(2 * j + 1));
        double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53; /* (0, 1] */
        double u2 = (double)(h2 >> 11) * 0x1.0p-53;         /* [0, 1) */
        double r = sqrt(-2.0 * log(u1));
        out[2 * j] = r * cos(two_pi * u2);
        if (2 * j + 1 < count) out[2 * j + 1] = r * sin(two_pi * u2);
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Brute-force DFT, the definition of DFT.hs:68-92 with the roots of Roots.hs:26-51.          */

/* out[k] = sum_j in[j] * exp(sign * 2 pi i * ((j*k) mod n) / n), in long double. */
int orc_dft_brute(int64_t n, int sign, const double *in, double *out)
{
    if (n < 0 || (sign != 1 && sign != -1)) return ORC_EINVAL;
    if (n == 0) return ORC_OK;
    long double *c = (long double *)malloc(sizeof(long double) * (size_t)n);
    long double *s = (long double *)malloc(sizeof(long double) * (size_t)n);
    if (!c || !s) { free(c); free(s); return ORC_ENOMEM; }
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < n; ++m) {
        long double t = two_pi * (long double)m / (long double)n;
        c[m] = cosl(t);
        s[m] = (long double)sign * sinl(t);
    }
    for (int64_t k = 0; k < n; ++k) {
        long double re = 0.0L, im = 0.0L;
        int64_t m = 0; /* (j*k) mod n, advanced exactly by k each step */
        for (int64_t j = 0; j < n; ++j) {
            long double xr = in[2 * j], xi = in[2 * j + 1];
            re += xr * c[m] - xi * s[m];
            im += xr * s[m] + xi * c[m];
            m += k;
            if (m >= n) m -= n;
        }
        out[2 * k] = (double)re;
        out[2 * k + 1] = (double)im;
    }
    free(c);
    free(s);
    return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Recursive mixed-radix decimation-in-time FFT, generic over the element type.               */

/* Twiddle table w[m] = exp(-2 pi i m / N) for m in [0, N), computed in long double. */
static long double *orc_twiddles_ld(int64_t N)
{
    long double *w = (long double *)malloc(sizeof(long double) * 2 * (size_t)(N > 0 ? N : 1));
    if (!w) return NULL;
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < N; ++m) {
        long double t = two_pi * (long double)m / (long double)N;
        w[2 * m] = cosl(t);
        w[2 * m + 1] = -sinl(t);
    }
    return w;
}

static int64_t orc_smallest_factor(int64_t n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int64_t p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/*
 * DEFINE_ORC_FFT(T, A, SUF): a 1D FFT in element type T with generic-prime accumulation in A.
 *   rec: out[0..n) = DFT_n of in[0], in[s], in[2s], ...  using the size-N twiddle table
 *        (tw index for W_n^e is e * (N/n)), sign folded in by conjugating the table reads.
 */
#define DEFINE_ORC_FFT(T, A, SUF)                                                               \
    static void orc_rec_##SUF(T *out, const T *in, int64_t n, int64_t s, const T *tw,          \
                              int64_t N, int sign, A *work)                                     \
    {                                                                                            \
        if (n == 1) {                                                                            \
            out[0] = in[0];                                                                      \
            out[1] = in[1];                                                                      \
            return;                                                                              \
        }                                                                                        \
        int64_t p = orc_smallest_factor(n);                                                      \
        int64_t m = n / p;                                                                       \
        int64_t step = N / n; /* W_n^e = tw[e * step] */                                        \
        for (int64_t r = 0; r < p; ++r)                                                          \
            orc_rec_##SUF(out + 2 * r * m, in + 2 * r * s, m, s * p, tw, N, sign, work);         \
        const T sg = (T)sign; /* forward table; conj for sign +1 */                             \
        if (p == 2) {                                                                            \
            for (int64_t k = 0; k < m; ++k) {                                                    \
                T *a = out + 2 * k, *b = out + 2 * (k + m);                                      \
                T wr = tw[2 * (k * step)];                                                       \
                T wi = (sign < 0) ? tw[2 * (k * step) + 1] : -tw[2 * (k * step) + 1];            \
                T br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;                        \
                T ar = a[0], ai = a[1];                                                          \
                a[0] = ar + br; a[1] = ai + bi;                                                  \
                b[0] = ar - br; b[1] = ai - bi;                                                  \
            }                                                                                    \
        } else if (p == 4) {                                                                     \
            for (int64_t k = 0; k < m; ++k) {                                                    \
                T vr[4], vi[4];                                                                  \
                for (int r = 0; r < 4; ++r) {                                                    \
                    const T *x = out + 2 * (k + r * m);                                          \
                    int64_t e = (int64_t)r * k * step;                                           \
                    T wr = tw[2 * e];                                                            \
                    T wi = (sign < 0) ? tw[2 * e + 1] : -tw[2 * e + 1];                          \
                    vr[r] = x[0] * wr - x[1] * wi;                                               \
                    vi[r] = x[0] * wi + x[1] * wr;                                               \
                }                                                                                \
                /* DFT_4 with j = sign * i: X_q = sum_r v_r (e^{sign 2pi i/4})^{rq} */          \
                T s0r = vr[0] + vr[2], s0i = vi[0] + vi[2];                                      \
                T d0r = vr[0] - vr[2], d0i = vi[0] - vi[2];                                      \
                T s1r = vr[1] + vr[3], s1i = vi[1] + vi[3];                                      \
                T d1r = vr[1] - vr[3], d1i = vi[1] - vi[3];                                      \
                /* (sign*i) * d1 = (-sign*d1i, sign*d1r) */                                      \
                T jd1r = -sg * d1i, jd1i = sg * d1r;                                             \
                T *y0 = out + 2 * k, *y1 = out + 2 * (k + m);                                    \
                T *y2 = out + 2 * (k + 2 * m), *y3 = out + 2 * (k + 3 * m);                      \
                y0[0] = s0r + s1r; y0[1] = s0i + s1i;                                            \
                y2[0] = s0r - s1r; y2[1] = s0i - s1i;                                            \
                y1[0] = d0r + jd1r; y1[1] = d0i + jd1i;                                          \
                y3[0] = d0r - jd1r; y3[1] = d0i - jd1i;                                          \
            }                                                                                    \
        } else {                                                                                 \
            /* generic prime p: X[k + q m] = sum_r W_n^{r k} W_p^{r q} y_r[k] */                 \
            A *v = work;          /* 2p */                                                       \
            A *acc = work + 2 * p; /* 2p */                                                      \
            for (int64_t k = 0; k < m; ++k) {                                                    \
                for (int64_t r = 0; r < p; ++r) {                                                \
                    const T *x = out + 2 * (k + r * m);                                          \
                    int64_t e = (r * k) % n;                                                     \
                    A wr = (A)tw[2 * (e * step)];                                                \
                    A wi = (A)((sign < 0) ? tw[2 * (e * step) + 1] : -tw[2 * (e * step) + 1]);   \
                    v[2 * r] = (A)x[0] * wr - (A)x[1] * wi;                                      \
                    v[2 * r + 1] = (A)x[0] * wi + (A)x[1] * wr;                                  \
                }                                                                                \
                for (int64_t q = 0; q < p; ++q) {                                                \
                    A sr = 0, si = 0;                                                            \
                    int64_t e = 0; /* (r*q) mod p, advanced exactly */                           \
                    for (int64_t r = 0; r < p; ++r) {                                            \
                        int64_t ee = e * m * step; /* W_p^e = W_N^{e N/p} */                     \
                        A wr = (A)tw[2 * ee];                                                    \
                        A wi = (A)((sign < 0) ? tw[2 * ee + 1] : -tw[2 * ee + 1]);               \
                        sr += v[2 * r] * wr - v[2 * r + 1] * wi;                                 \
                        si += v[2 * r] * wi + v[2 * r + 1] * wr;                                 \
                        e += q;                                                                  \
                        if (e >= p) e -= p;                                                      \
                    }                                                                            \
                    acc[2 * q] = sr;                                                             \
                    acc[2 * q + 1] = si;                                                         \
                }                                                                                \
                for (int64_t q = 0; q < p; ++q) {                                                \
                    out[2 * (k + q * m)] = (T)acc[2 * q];                                        \
                    out[2 * (k + q * m) + 1] = (T)acc[2 * q + 1];                                \
                }                                                                                \
            }                                                                                    \
        }                                                                                        \
    }                                                                                            \
                                                                                                 \
    static int64_t orc_max_prime_##SUF(int64_t n)                                                \
    {                                                                                            \
        int64_t best = 1;                                                                        \
        while (n > 1) {                                                                          \
            int64_t p = orc_smallest_factor(n);                                                  \
            if (p > best) best = p;                                                              \
            n /= p;                                                                              \
        }                                                                                        \
        return best;                                                                             \
    }                                                                                            \
                                                                                                 \
    /* howmany contiguous rows of length n, unnormalised, sign -1/+1, then times scale. */       \
    int orc_fft_rows_##SUF(int64_t n, int sign, int64_t howmany, double scale, const T *in,      \
                           T *out, int nthreads)                                                 \
    {                                                                                            \
        if (n < 0 || howmany < 0 || (sign != 1 && sign != -1)) return ORC_EINVAL;                \
        if (n == 0 || howmany == 0) return ORC_OK;                                               \
        long double *wl = orc_twiddles_ld(n);                                                    \
        T *tw = (T *)malloc(sizeof(T) * 2 * (size_t)n);                                          \
        if (!wl || !tw) { free(wl); free(tw); return ORC_ENOMEM; }                               \
        for (int64_t i = 0; i < 2 * n; ++i) tw[i] = (T)wl[i];                                    \
        free(wl);                                                                                \
        int64_t pmax = orc_max_prime_##SUF(n);                                                   \
        int err = ORC_OK;                                                                        \
        if (nthreads < 1) nthreads = 1;                                                          \
        _Pragma("omp parallel num_threads(nthreads)")                                            \
        {                                                                                        \
            A *work = (A *)malloc(sizeof(A) * 4 * (size_t)(pmax + 1));                           \
            if (!work) {                                                                         \
                _Pragma("omp atomic write") err = ORC_ENOMEM;                                    \
            } else {                                                                             \
                _Pragma("omp for schedule(static)")                                              \
                for (int64_t b = 0; b < howmany; ++b) {                                          \
                    orc_rec_##SUF(out + 2 * b * n, in + 2 * b * n, n, 1, tw, n, sign, work);     \
                    if (scale != 1.0) {                                                          \
                        T sc = (T)scale;                                                         \
                        for (int64_t i = 0; i < 2 * n; ++i) out[2 * b * n + i] *= sc;            \
                    }                                                                            \
                }                                                                                \
                free(work);                                                                      \
            }                                                                                    \
        }                                                                                        \
        free(tw);                                                                                \
        return err;                                                                              \
    }

DEFINE_ORC_FFT(double, long double, f64)
DEFINE_ORC_FFT(float, double, f32)

/* ------------------------------------------------------------------------------------------ */
/* N-dimensional: the 1D transform applied along every axis whose bit is set in axes_mask.    */
/* Axis a counts from the outermost (a = 0) to the innermost (a = rank-1), the order of the   */
/* Accelerate shape Z :. extent_0 :. ... :. extent_{rank-1}.                                   */

int orc_fft_nd_f64(int rank, const int64_t *extents, int axes_mask, int sign, double scale,
                   const double *in, double *out, int nthreads)
{
    if (rank < 1 || rank > 8 || (sign != 1 && sign != -1)) return ORC_EINVAL;
    int64_t total = 1;
    for (int a = 0; a < rank; ++a) {
        if (extents[a] < 0) return ORC_EINVAL;
        total *= extents[a];
    }
    if (total == 0) return ORC_OK;
    memmove(out, in, sizeof(double) * 2 * (size_t)total);
    if (nthreads < 1) nthreads = 1;
    for (int a = rank - 1; a >= 0; --a) {
        if (!(axes_mask & (1 << a))) continue;
        int64_t n = extents[a];
        if (n == 1) continue;
        int64_t inner = 1;
        for (int b = a + 1; b < rank; ++b) inner *= extents[b];
        int64_t outer = total / (n * inner);
        if (inner == 1) {
            /* the recursion is out-of-place: transform from a copy */
            double *copy = (double *)malloc(sizeof(double) * 2 * (size_t)total);
            if (!copy) return ORC_ENOMEM;
            memcpy(copy, out, sizeof(double) * 2 * (size_t)total);
            int err = orc_fft_rows_f64(n, sign, outer, 1.0, copy, out, nthreads);
            free(copy);
            if (err) return err;
            continue;
        }
        /* gather each strided line into a contiguous scratch line, transform, scatter back */
        double *line = (double *)malloc(sizeof(double) * 4 * (size_t)n * (size_t)inner);
        if (!line) return ORC_ENOMEM;
        for (int64_t o = 0; o < outer; ++o) {
            double *base = out + 2 * o * n * inner;
            double *src = line, *dst = line + 2 * n * inner;
            for (int64_t c = 0; c < inner; ++c)
                for (int64_t i = 0; i < n; ++i) {
                    src[2 * (c * n + i)] = base[2 * (i * inner + c)];
                    src[2 * (c * n + i) + 1] = base[2 * (i * inner + c) + 1];
                }
            int err = orc_fft_rows_f64(n, sign, inner, 1.0, src, dst, nthreads);
            if (err) { free(line); return err; }
            for (int64_t c = 0; c < inner; ++c)
                for (int64_t i = 0; i < n; ++i) {
                    base[2 * (i * inner + c)] = dst[2 * (c * n + i)];
                    base[2 * (i * inner + c) + 1] = dst[2 * (c * n + i) + 1];
                }
        }
        free(line);
    }
    if (scale != 1.0)
        for (int64_t i = 0; i < 2 * total; ++i) out[i] *= scale;
    return ORC_OK;
}
