"""Writes tests/golden/dft_mpmath.npz: known-answer DFT vectors computed with mpmath at 40
significant digits, independent of every FFT in this repository.

The reference (accelerate-fft v1.3.0.0) holds no golden vectors, known-answer tests or fixtures
(its tests are the seven properties of test/Test/FFT.hs:112-236), so these vectors pin absolute
output values from the DFT definition itself: src/Data/Array/Accelerate/Math/DFT.hs:68-92 with the
roots of DFT/Roots.hs:26-51, Forward = exp(-2 pi i jk/n), Reverse = exp(+2 pi i jk/n) unnormalised
(FFT/Mode.hs:15-26).

Inputs are re/im uniform[-1,1) from a counter-based SplitMix64 stream restated here in pure Python
(re of element i from counter 2i, im from 2i+1), so the fixture also pins the C generator used for
seeded inputs.  Sizes: the generator range of test/Test/Base.hs:44-45 (1..1024) sampled at primes,
5-smooth and odd sizes, plus the full N=1024 seed-0 vector pair of the Complex Double parity case.

Run:  python tests/golden/make_golden.py      (about a minute; mpmath is pure Python)
"""
import os

import mpmath
import numpy as np

M64 = (1 << 64) - 1


def mix64(z):
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def uniform_stream(seed, count):
    base = mix64(seed)
    out = []
    for i in range(count):
        h = mix64((base + i) & M64)
        u = (h >> 11) * 2.0 ** -53
        out.append(2.0 * u - 1.0)
    return out


def dft(xs, sign):
    n = len(xs)
    roots = [mpmath.mpc(mpmath.cospi(mpmath.mpf(2 * m) / n), sign * mpmath.sinpi(mpmath.mpf(2 * m) / n))
             for m in range(n)]
    xm = [mpmath.mpc(re, im) for re, im in xs]
    out = []
    for k in range(n):
        acc = mpmath.mpc(0)
        m = 0
        for j in range(n):
            acc += xm[j] * roots[m]
            m += k
            if m >= n:
                m -= n
        out.append(complex(acc))
    return np.array(out, dtype=np.complex128)


def main():
    mpmath.mp.dps = 40
    sizes = [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 15, 16, 17, 25, 27, 30, 31, 32, 48, 60, 64,
             97, 100, 120, 127, 128, 243, 256, 1024]
    arrays = {}
    for n in sizes:
        seed = 0 if n == 1024 else 1000 + n
        flat = uniform_stream(seed, 2 * n)
        xs = [(flat[2 * i], flat[2 * i + 1]) for i in range(n)]
        x = np.array([complex(a, b) for a, b in xs], dtype=np.complex128)
        arrays[f"x_{n}"] = x
        arrays[f"seed_{n}"] = np.array([seed], dtype=np.int64)
        arrays[f"fwd_{n}"] = dft(xs, -1)
        arrays[f"rev_{n}"] = dft(xs, +1)
        print("n =", n, flush=True)
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "dft_mpmath.npz"), **arrays)


if __name__ == "__main__":
    main()
