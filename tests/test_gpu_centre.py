"""DFT/Centre.hs on the GPU: the standalone centre / shift / ishift (bit-exact against the oracle)
and their fused form around fft1D/2D/3D (afft_fft_centred), whose signs ride on the first load or
the last store of every kind of pass the planner emits."""
import itertools

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode, centre as C
from conftest import tolerance

pytestmark = pytest.mark.gpu

DTYPES = [torch.complex64, torch.complex128]
OMODE = {Mode.Forward: O.Mode.Forward, Mode.Reverse: O.Mode.Reverse, Mode.Inverse: O.Mode.Inverse}
OPS = {None: lambda a: a, "centre": O.centre, "shift": O.shift, "ishift": O.ishift}
GPU_OPS = {1: (C.centre1D, C.shift1D, C.ishift1D), 2: (C.centre2D, C.shift2D, C.ishift2D),
           3: (C.centre3D, C.shift3D, C.ishift3D)}


def dev(x, dtype, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).to(device)


def host(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def fftn(mode, a):
    return {1: O.fft1D, 2: O.fft2D, 3: O.fft3D}[a.ndim](OMODE[mode], a)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(8,), (7,), (1,), (4, 6), (5, 6), (3, 4, 5), (4, 4, 8), (1, 3, 1)])
def test_reorder_bit_exact(gpu, dtype, shape):
    x = dev(O.uniform(3, shape), dtype, gpu)
    xr = host(x)
    centre, shift, ishift = GPU_OPS[len(shape)]
    np.testing.assert_array_equal(host(centre(x)), O.centre(xr))
    np.testing.assert_array_equal(host(shift(x)), O.shift(xr))
    np.testing.assert_array_equal(host(ishift(x)), O.ishift(xr))
    np.testing.assert_array_equal(host(ishift(shift(x))), xr)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(64,), (63,), (32, 48), (31, 48), (8, 6, 10), (5, 6, 7)])
def test_fft_centred_every_combination(gpu, dtype, shape):
    x = dev(O.normal(9, shape), dtype, gpu)
    xr = host(x)
    n = int(np.prod(shape))
    tol = tolerance(dtype == torch.complex128, n)
    for pre, post in itertools.product(OPS, OPS):
        for mode in (Mode.Forward, Mode.Inverse):
            want = OPS[post](fftn(mode, OPS[pre](xr)))
            got = C.fft_centred(mode, x, pre=pre, post=post)
            err = O.rel_l2(host(got), want)
            assert err <= tol, f"{shape} pre={pre} post={post} {mode}: {err:.3e}"


def test_reference_identity_on_the_gpu(gpu):
    # fft(centre(X)) = shift(fft(X)) (Centre.hs:17-19), both sides through the engine
    x = dev(O.normal(4, (64, 96)), torch.complex128, gpu)
    lhs = C.fft_centred(Mode.Forward, x, pre="centre")
    rhs = C.shift2D(afft.fft2D(Mode.Forward, x))
    assert O.rel_l2(host(lhs), host(rhs)) <= 1e-13


@pytest.mark.parametrize("dtype,shape", [
    (torch.complex64, (4096, 4096)),    # row kernel first, register-resident column pass B last
    (torch.complex64, (20480,)),        # six-step: register column pass first, transpose last
    (torch.complex64, (257,)),          # fused Bluestein
    (torch.complex64, (10007,)),        # multi-kernel Bluestein: chirp pre first, chirp post last
    (torch.complex128, (40960,)),       # balanced four-step
    (torch.complex128, (64, 48, 80)),   # rows then strided LDS passes
    (torch.complex64, (2048, 64)),      # split strided columns
    (torch.complex64, (1 << 20,)),      # balanced four-step of 1024 x 1024
])
def test_fused_signs_on_every_pass_kind(gpu, dtype, shape):
    """The fused forms against the unfused composition of standalone reorders and plain
    transforms (both pinned at small sizes above), at sizes that select each plan shape."""
    x = dev(O.uniform(8, shape), dtype, gpu)
    rank = len(shape)
    centre, shift, ishift = GPU_OPS[rank]
    plain = {1: afft.fft1D, 2: afft.fft2D, 3: afft.fft3D}[rank]
    tol = tolerance(dtype == torch.complex128, int(np.prod(shape)))
    even = all(s % 2 == 0 for s in shape)
    cases = [("centre", None, lambda: plain(Mode.Forward, centre(x))),
             (None, "centre", lambda: centre(plain(Mode.Forward, x))),
             (None, "shift", lambda: shift(plain(Mode.Forward, x))),
             ("ishift", "centre", lambda: centre(plain(Mode.Forward, ishift(x))))]
    for pre, post, ref in cases:
        got = C.fft_centred(Mode.Forward, x, pre=pre, post=post)
        err = O.rel_l2(host(got), host(ref()))
        assert err <= tol, f"{shape} pre={pre} post={post} (even={even}): {err:.3e}"


@pytest.mark.parametrize("shape", [(32, 48), (31, 48), (4096,)])
def test_fft_centred_in_place(gpu, shape):
    x = dev(O.uniform(2, shape), torch.complex64, gpu)
    xr = host(x)
    want = O.shift(fftn(Mode.Inverse, O.centre(xr)))
    y = x.clone()
    C.fft_centred(Mode.Inverse, y, pre="centre", post="shift", out=y)
    assert O.rel_l2(host(y), want) <= tolerance(False, int(np.prod(shape)))


def test_reorder_rejects():
    lib = afft._lib.lib()
    e = afft._lib.extents_arg([4, 4, 4, 4])
    assert lib.afft_reorder(1, 0, 4, e, None, None, None) == 2          # rank 4: unsupported rank
    assert lib.afft_reorder(9, 0, 1, e, None, None, None) == 1          # unknown operation
    assert lib.afft_fft_centred(0, 0, 2, e, 1, 7, None, None, None) == 1
    big = afft._lib.extents_arg([1 << 16, 1 << 15])
    assert lib.afft_reorder(1, 0, 2, big, None, None, None) == 4        # 2^31 elements: too large
