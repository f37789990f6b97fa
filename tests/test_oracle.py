"""The CPU oracle, pinned before anything is checked against it.

The reference holds no golden vectors (its tests are the seven properties of
test/Test/FFT.hs:112-236), so the oracle is pinned by
  * mpmath 40-digit DFT fixtures (tests/golden/dft_mpmath.npz, made by make_golden.py),
  * the long-double brute-force definition (DFT.hs:68-92),
  * closed forms (impulse, constant, pure tone) that fix the sign convention of Mode.hs:21-26,
  * the seven reference properties, and
  * planted mistakes (wrong sign, dropped twiddle, transposed axes) that the checks must catch.
"""
import math
import os

import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle as O

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dft_mpmath.npz")


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def sizes(g):
    return sorted(int(k[2:]) for k in g if k.startswith("x_"))


def test_generator_matches_fixture_inputs(golden):
    # the C SplitMix64 stream equals the pure-Python restatement in make_golden.py
    for n in sizes(golden):
        x = O.uniform(int(golden[f"seed_{n}"][0]), (n,))
        np.testing.assert_array_equal(x, golden[f"x_{n}"])


def test_generator_ranges():
    u = O.uniform(7, (100000,))
    assert u.real.min() >= -1 and u.real.max() < 1 and u.imag.min() >= -1 and u.imag.max() < 1
    assert abs(u.real.mean()) < 0.01 and abs(u.real.var() - 1 / 3) < 0.01
    z = O.normal(7, (100000,))
    assert abs(z.real.mean()) < 0.02 and abs(z.real.std() - 1) < 0.02
    assert abs(z.imag.std() - 1) < 0.02
    # float rounding happens after generation: complex64 values equal rounded complex128 ones
    np.testing.assert_array_equal(O.uniform(3, (64,), np.complex64), O.uniform(3, (64,)).astype(np.complex64))


def test_brute_force_matches_mpmath(golden):
    for n in sizes(golden):
        x = golden[f"x_{n}"]
        assert O.rel_l2(O.dft_brute(x, -1), golden[f"fwd_{n}"]) < 1e-16 * max(4, n)
        assert O.rel_l2(O.dft_brute(x, +1), golden[f"rev_{n}"]) < 1e-16 * max(4, n)


def test_fft_matches_mpmath(golden):
    for n in sizes(golden):
        x = golden[f"x_{n}"]
        assert O.rel_l2(O.fft_rows(x, -1), golden[f"fwd_{n}"]) < 1e-15, n
        assert O.rel_l2(O.fft_rows(x, +1), golden[f"rev_{n}"]) < 1e-15, n


@pytest.mark.parametrize("n", list(range(1, 70)) + [97, 121, 125, 128, 243, 255, 256, 257, 1000, 1009, 1024])
def test_fft_matches_brute_force(n):
    x = O.uniform(n, (n,))
    for sign in (-1, 1):
        assert O.rel_l2(O.fft_rows(x, sign), O.dft_brute(x, sign)) < 5e-16 * max(1.0, math.log2(n)) + 1e-16


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 8, 12, 17, 64, 97, 100, 1024, 4099])
def test_closed_forms(n):
    k = np.arange(n)
    # impulse at n0 -> exp(-2 pi i k n0 / N) under Forward, exp(+...) under Reverse
    n0 = (3 * n) // 7
    d = np.zeros(n, np.complex128)
    d[n0] = 1
    want = np.exp(-2j * np.pi * ((k * n0) % n) / n)
    assert O.rel_l2(O.fft1D(O.Mode.Forward, d), want) < 1e-14
    assert O.rel_l2(O.fft1D(O.Mode.Reverse, d), np.conj(want)) < 1e-14
    assert O.rel_l2(O.fft1D(O.Mode.Inverse, d), np.conj(want) / n) < 1e-14
    # constant c -> N c delta[k]
    c = np.full(n, 0.25 - 0.5j)
    want = np.zeros(n, np.complex128)
    want[0] = n * (0.25 - 0.5j)
    assert O.rel_l2(O.fft1D(O.Mode.Forward, c), want) < 1e-14
    # tone exp(+2 pi i k0 n / N) -> N delta[k - k0] under Forward
    k0 = (2 * n) // 5
    tone = np.exp(2j * np.pi * ((k * k0) % n) / n)
    want = np.zeros(n, np.complex128)
    want[k0] = n
    assert O.rel_l2(O.fft1D(O.Mode.Forward, tone), want) < 1e-13


def test_mode_semantics_and_scale():
    # FFT.hs:63-84 (innermost extent), :92-111 (length), :119-173 (total size)
    x = O.uniform(11, (6, 10))
    fwd = O.fft(O.Mode.Forward, x)
    np.testing.assert_allclose(O.fft(O.Mode.Inverse, fwd), x, atol=1e-14)
    np.testing.assert_allclose(O.fft(O.Mode.Reverse, fwd), 10 * x, atol=1e-13)
    for i in range(6):  # innermost only, rows independent
        np.testing.assert_allclose(fwd[i], O.dft_brute(x[i], -1), atol=1e-13)
    f2 = O.fft2D(O.Mode.Forward, x)
    np.testing.assert_allclose(O.fft2D(O.Mode.Inverse, f2), x, atol=1e-14)
    np.testing.assert_allclose(O.fft2D(O.Mode.Reverse, f2), 60 * x, atol=1e-12)
    y = O.uniform(12, (3, 4, 5))
    f3 = O.fft3D(O.Mode.Forward, y)
    np.testing.assert_allclose(O.fft3D(O.Mode.Inverse, f3), y, atol=1e-14)


def test_separability_against_brute_force():
    # 2D = column DFT of row DFTs, each by the O(N^2) definition
    x = O.uniform(13, (5, 12))
    rows = np.stack([O.dft_brute(r, -1) for r in x])
    want = np.stack([O.dft_brute(rows[:, j], -1) for j in range(12)], axis=1)
    assert O.rel_l2(O.fft2D(O.Mode.Forward, x), want) < 1e-15
    y = O.uniform(14, (3, 4, 6))
    w = y.copy()
    for axis in (2, 1, 0):
        w = np.apply_along_axis(lambda v: O.dft_brute(v, +1), axis, w)
    assert O.rel_l2(O.fft_axes(y, (0, 1, 2), +1), w) < 1e-15


def test_planted_mistakes_are_caught():
    # The checks above must fail on plausible mistakes planted in the oracle's output.
    n = 12
    k = np.arange(n)
    d = np.zeros(n, np.complex128)
    d[5] = 1
    want = np.exp(-2j * np.pi * ((k * 5) % n) / n)
    wrong_sign = O.fft_rows(d, +1)
    assert O.rel_l2(wrong_sign, want) > 0.1
    # a dropped twiddle: radix-2 combine without W_n^k
    x = O.uniform(15, (n,))
    e, o = O.fft_rows(x[0::2], -1), O.fft_rows(x[1::2], -1)
    dropped = np.concatenate([e + o, e - o])
    assert O.rel_l2(dropped, O.dft_brute(x, -1)) > 0.1
    # the wrong axis in fft: the outer axis transformed instead of the innermost one
    x2 = O.uniform(16, (4, 6))
    rows = np.stack([O.dft_brute(r, -1) for r in x2])
    wrong_axis = O.fft_axes(x2, (0,), -1)
    assert O.rel_l2(wrong_axis, rows) > 0.1
    ref = O.fft2D(O.Mode.Forward, x2)
    # a wrong Inverse divisor (innermost extent instead of total size in fft2D)
    bad = O.fft_axes(ref, (0, 1), +1, 1.0 / 6)
    assert O.rel_l2(bad, x2) > 0.1


# --- the seven properties of test/Test/FFT.hs:112-236 on the oracle ------------------------
# generators of test/Test/Base.hs:35-61: values uniform in [-1, 1], DIM1 n in [1, 1024],
# DIM2 1..48 x 1..128, DIM3 1..16 x 1..32 x 1..64

modes = st.sampled_from([O.Mode.Forward, O.Mode.Reverse, O.Mode.Inverse])
dim1 = st.tuples(st.integers(1, 1024))
dim2 = st.tuples(st.integers(1, 48), st.integers(1, 128))
dim3 = st.tuples(st.integers(1, 16), st.integers(1, 32), st.integers(1, 64))
TRANSFORMS = {1: O.fft1D, 2: O.fft2D, 3: O.fft3D}


def rev(a):
    n = a.shape[-1]
    return a[..., (-np.arange(n)) % n]


def similar(a, b, tol=1e-12):
    return O.rel_l2(a, b) <= tol or np.max(np.abs(a - b)) <= tol


@settings(max_examples=25, deadline=None)
@given(mode=modes, sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31), c=st.complex_numbers(max_magnitude=1))
def test_homogeneity(mode, sh, seed, c):
    x = O.uniform(seed, sh)
    T = TRANSFORMS[len(sh)]
    assert similar(T(mode, c * x), c * T(mode, x))


@settings(max_examples=25, deadline=None)
@given(mode=modes, sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_additivity(mode, sh, seed):
    x, y = O.uniform(seed, sh), O.uniform(seed + 1, sh)
    T = TRANSFORMS[len(sh)]
    assert similar(T(mode, x + y), T(mode, x) + T(mode, y))


@settings(max_examples=25, deadline=None)
@given(sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_inverse(sh, seed):
    x = O.uniform(seed, sh)
    T = TRANSFORMS[len(sh)]
    assert similar(T(O.Mode.Inverse, T(O.Mode.Forward, x)), x)


@settings(max_examples=25, deadline=None)
@given(mode=modes, sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_reverse(mode, sh, seed):
    x = O.uniform(seed, sh)
    assert similar(rev(O.fft(mode, x)), O.fft(mode, rev(x)))


@settings(max_examples=25, deadline=None)
@given(mode=modes, sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_conjugate(mode, sh, seed):
    x = O.uniform(seed, sh)
    assert similar(np.conj(O.fft(mode, x)), O.fft(mode, np.conj(rev(x))))


@settings(max_examples=25, deadline=None)
@given(mode=st.sampled_from([O.Mode.Forward, O.Mode.Reverse]), sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_isometry(mode, sh, seed):
    x = O.uniform(seed, sh)
    n = sh[-1]
    got = np.sqrt(np.sum(np.abs(O.fft(mode, x)) ** 2, axis=-1))
    want = math.sqrt(n) * np.sqrt(np.sum(np.abs(x) ** 2, axis=-1))
    np.testing.assert_allclose(got, want, rtol=1e-12)


@settings(max_examples=25, deadline=None)
@given(mode=st.sampled_from([O.Mode.Forward, O.Mode.Reverse]), sh=st.one_of(dim1, dim2, dim3), seed=st.integers(0, 2**31))
def test_unitarity(mode, sh, seed):
    x, y = O.uniform(seed, sh), O.uniform(seed + 7, sh)
    n = sh[-1]
    got = np.sum(O.fft(mode, x) * np.conj(O.fft(mode, y)), axis=-1)
    want = n * np.sum(x * np.conj(y), axis=-1)
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11 * n)


def test_large_sizes_against_numpy_pocketfft():
    # third-party cross-check at the large sizes (pocketfft, FP64): not the reference, but an
    # independent implementation of the same definition
    x = O.uniform(2, (8, 4096))
    assert O.rel_l2(O.fft(O.Mode.Forward, x), np.fft.fft(x, axis=-1)) < 1e-15
    x = O.uniform(3, (1 << 22,))
    assert O.rel_l2(O.fft1D(O.Mode.Forward, x), np.fft.fft(x)) < 1e-15
    y = O.normal(4, (64, 48, 40))
    assert O.rel_l2(O.fft3D(O.Mode.Inverse, y), np.fft.ifftn(y)) < 1e-15


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1000, 4096, 4099])
def test_float_variant_accuracy(n):
    # the float recursion (used only to time a CPU path beside the GPU) stays within the FP32
    # tolerance of the FP64 oracle
    x = O.uniform(n + 5, (16, n))
    xf = x.astype(np.complex64)
    want = O.fft_rows(xf.astype(np.complex128), -1)
    got = O.fft_rows_f32(xf, -1)
    assert O.rel_l2(got, want) <= 1e-6 * max(1.0, math.log2(max(n, 2)))


# ----------------------------------------------------------------------------------------------
# DFT/Centre.hs restatement

@pytest.mark.parametrize("shape", [(1,), (5,), (8,), (3, 4), (4, 5), (2, 3, 5), (4, 4, 2)])
def test_centre_shift_against_numpy_and_identities(shape):
    x = O.uniform(31, shape)
    # shift is numpy's fftshift and ishift its inverse (an independent library routine)
    np.testing.assert_array_equal(O.shift(x), np.fft.fftshift(x))
    np.testing.assert_array_equal(O.ishift(x), np.fft.ifftshift(x))
    # ishift . shift = id (Centre.hs:84-86)
    np.testing.assert_array_equal(O.ishift(O.shift(x)), x)
    # centre is an involution with the sign of the index parity
    np.testing.assert_array_equal(O.centre(O.centre(x)), x)
    signs = np.indices(shape).sum(axis=0) % 2
    np.testing.assert_array_equal(O.centre(x), np.where(signs == 1, -x, x))
    # fft(centre(X)) = shift(fft(X)) (Centre.hs:17-19) when every extent is even
    F = O.fft_axes(x, tuple(range(len(shape))), -1)
    if all(n % 2 == 0 for n in shape):
        np.testing.assert_allclose(O.fft_axes(O.centre(x), tuple(range(len(shape))), -1), O.shift(F),
                                   rtol=0, atol=1e-12 * np.abs(F).max())


def test_shift_known_answers():
    v = np.arange(5, dtype=np.complex128)
    np.testing.assert_array_equal(O.shift(v).real, [3, 4, 0, 1, 2])   # roll by ceil(5/2) = 3
    np.testing.assert_array_equal(O.ishift(v).real, [2, 3, 4, 0, 1])  # roll by floor(5/2) = 2
    np.testing.assert_array_equal(O.centre(v).real, [0, -1, 2, -3, 4])


def test_planted_centre_mistakes_fail():
    x = O.uniform(5, (5, 6))
    # y[i] = x[(i + s) mod n] is np.roll by -s; the right shift rolls the odd axis by ceil(5/2) = 3
    np.testing.assert_array_equal(O.shift(x), np.roll(x, (-3, -3), axis=(0, 1)))
    wrong_shift = np.roll(x, (-2, -3), axis=(0, 1))    # floor instead of ceil on the odd axis
    assert not np.array_equal(O.shift(x), wrong_shift)
    wrong_centre = x * (-1.0) ** np.indices(x.shape)[1]  # only the innermost index
    assert not np.array_equal(O.centre(x), wrong_centre)


# ----------------------------------------------------------------------------------------------
# Math/DFT.hs restatement

def _std_roots(shape, sign):
    n = shape[-1]
    k = 2 * np.pi * np.arange(n) / n
    return np.broadcast_to(np.cos(k) + 1j * sign * np.sin(k), shape)


@pytest.mark.parametrize("n", [1, 2, 5, 16, 31, 64])
def test_dftG_vector_is_the_dft(n):
    x = O.uniform(n, (n,))
    for sign in (-1, 1):
        got = O.dftG(_std_roots((n,), sign), x)
        np.testing.assert_allclose(got, O.dft_brute(x, sign), rtol=0, atol=1e-12 * max(1, n))


def test_dftG_higher_rank_sums_every_row():
    # the reference folds over shapeSize elements (DFT.hs:79-92): every row gets the DFT of the
    # row sum when all rows share the roots
    x = O.uniform(3, (4, 12))
    got = O.dftG(_std_roots(x.shape, -1), x)
    want = O.dft_brute(x.sum(axis=0), -1)
    for r in range(4):
        np.testing.assert_allclose(got[r], want, rtol=0, atol=1e-12)
    # and per-row roots weight each row by its own roots
    roots = _std_roots(x.shape, -1) * np.arange(1, 5)[:, None]
    got2 = O.dftG(roots, x)
    want2 = sum(O.dft_brute(x[r] * (r + 1), -1) for r in range(4))
    np.testing.assert_allclose(got2[0], want2, rtol=0, atol=1e-11)
