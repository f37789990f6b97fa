"""The seven properties of test/Test/FFT.hs:112-236 through the GPU path, with the generators of
test/Test/Base.hs:35-61 (values uniform in [-1, 1], DIM1 n in [1, 1024], DIM2 1..48 x 1..128,
DIM3 1..16 x 1..32 x 1..64), for Complex Float and Complex Double."""
import math

import numpy as np
import pytest
import torch
from hypothesis import given, settings, strategies as st, HealthCheck

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu

modes = st.sampled_from([Mode.Forward, Mode.Reverse, Mode.Inverse])
dim1 = st.tuples(st.integers(1, 1024))
dim2 = st.tuples(st.integers(1, 48), st.integers(1, 128))
dim3 = st.tuples(st.integers(1, 16), st.integers(1, 32), st.integers(1, 64))
shapes = st.one_of(dim1, dim2, dim3)
dtypes = st.sampled_from([torch.complex64, torch.complex128])
TRANSFORM = {1: afft.fft1D, 2: afft.fft2D, 3: afft.fft3D}
SETTINGS = settings(max_examples=30, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dtype).to("cuda")


def host(t):
    return t.cpu().numpy().astype(np.complex128)


def similar(a, b, dtype, n):
    # the reference compares with ~~~ (accelerate's Similar); here: relative L2 at the stated
    # tolerance, scaled by a small factor for the extra rounding of composed transforms
    tol = 4 * tolerance(dtype == torch.complex128, n)
    return O.rel_l2(a, b) <= tol or np.max(np.abs(a - b)) <= tol


def rev(a):
    n = a.shape[-1]
    return a[..., (-np.arange(n)) % n]


@SETTINGS
@given(mode=modes, sh=shapes, seed=st.integers(0, 2**31), c=st.complex_numbers(max_magnitude=1), dtype=dtypes)
def test_homogeneity(gpu, mode, sh, seed, c, dtype):
    x = O.uniform(seed, sh)
    T = TRANSFORM[len(sh)]
    xd = dev(x, dtype)
    cd = torch.tensor(c, dtype=dtype, device="cuda")
    assert similar(host(T(mode, cd * xd)), host(cd * T(mode, xd)), dtype, int(np.prod(sh)))


@SETTINGS
@given(mode=modes, sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_additivity(gpu, mode, sh, seed, dtype):
    T = TRANSFORM[len(sh)]
    x, y = dev(O.uniform(seed, sh), dtype), dev(O.uniform(seed + 1, sh), dtype)
    assert similar(host(T(mode, x + y)), host(T(mode, x) + T(mode, y)), dtype, int(np.prod(sh)))


@SETTINGS
@given(sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_inverse(gpu, sh, seed, dtype):
    T = TRANSFORM[len(sh)]
    x = dev(O.uniform(seed, sh), dtype)
    assert similar(host(T(Mode.Inverse, T(Mode.Forward, x))), host(x), dtype, int(np.prod(sh)))


@SETTINGS
@given(mode=modes, sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_reverse(gpu, mode, sh, seed, dtype):
    x = O.uniform(seed, sh)
    got = rev(host(afft.fft(mode, dev(x, dtype))))
    want = host(afft.fft(mode, dev(rev(x), dtype)))
    assert similar(got, want, dtype, sh[-1])


@SETTINGS
@given(mode=modes, sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_conjugate(gpu, mode, sh, seed, dtype):
    x = O.uniform(seed, sh)
    got = np.conj(host(afft.fft(mode, dev(x, dtype))))
    want = host(afft.fft(mode, dev(np.conj(rev(x)), dtype)))
    assert similar(got, want, dtype, sh[-1])


@SETTINGS
@given(mode=st.sampled_from([Mode.Forward, Mode.Reverse]), sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_isometry(gpu, mode, sh, seed, dtype):
    x = dev(O.uniform(seed, sh), dtype)
    n = sh[-1]
    got = host(torch.sqrt(torch.sum(torch.abs(afft.fft(mode, x)) ** 2, dim=-1)))
    want = math.sqrt(n) * host(torch.sqrt(torch.sum(torch.abs(x) ** 2, dim=-1)))
    assert similar(got, want, dtype, n)


@SETTINGS
@given(mode=st.sampled_from([Mode.Forward, Mode.Reverse]), sh=shapes, seed=st.integers(0, 2**31), dtype=dtypes)
def test_unitarity(gpu, mode, sh, seed, dtype):
    x, y = dev(O.uniform(seed, sh), dtype), dev(O.uniform(seed + 7, sh), dtype)
    n = sh[-1]
    got = host(torch.sum(afft.fft(mode, x) * torch.conj(afft.fft(mode, y)), dim=-1))
    want = n * host(torch.sum(x * torch.conj(y), dim=-1))
    assert similar(got, want, dtype, n)
