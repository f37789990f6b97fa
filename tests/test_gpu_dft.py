"""Math.DFT on the GPU (afft_dft / afft_dftG / afft_dftGS) against the oracle's restatement of
DFT.hs:63-111 and against the brute-force DFT; tolerances as for the FFT."""
import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import dft as D
from conftest import tolerance

pytestmark = pytest.mark.gpu

DTYPES = [torch.complex64, torch.complex128]


def dev(x, dtype, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).to(device)


def host(t):
    return t.detach().cpu().numpy().astype(np.complex128)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [1, 7, 64, 100, 257, 1024])
def test_dft_idft_vector(gpu, dtype, n):
    x = dev(O.uniform(n, (n,)), dtype, gpu)
    xr = host(x)
    tol = tolerance(dtype == torch.complex128, n)
    assert O.rel_l2(host(D.dft(x)), O.dft_brute(xr, -1)) <= tol
    assert O.rel_l2(host(D.idft(x)), O.dft_brute(xr, +1) / n) <= tol
    # agrees with the FFT path on the same vector
    assert O.rel_l2(host(D.dft(x)), host(afft.fft1D(afft.Mode.Forward, x))) <= 2 * tol


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(3, 16), (2, 3, 10), (5, 1)])
def test_dftG_reference_semantics(gpu, dtype, shape):
    x = dev(O.uniform(2, shape), dtype, gpu)
    roots = D.rootsOfUnity(shape, dtype, gpu)
    got = host(D.dftG(roots, x))
    want = O.dftG(host(roots), host(x))
    n = int(np.prod(shape))
    assert O.rel_l2(got, want) <= tolerance(dtype == torch.complex128, n)
    # every row holds the same values (the fold runs over all elements)
    flat = got.reshape(-1, shape[-1])
    assert np.array_equal(flat, np.broadcast_to(flat[0], flat.shape))


@pytest.mark.parametrize("dtype", DTYPES)
def test_dftG_custom_roots_and_dftGS(gpu, dtype):
    shape = (4, 12)
    x = dev(O.uniform(4, shape), dtype, gpu)
    roots = dev(O.uniform(5, shape), dtype, gpu)  # arbitrary roots of the same extent
    want = O.dftG(host(roots), host(x))
    tol = tolerance(dtype == torch.complex128, 48)
    assert O.rel_l2(host(D.dftG(roots, x)), want) <= tol
    for ix in [(0, 0), (3, 5), (1, 11)]:
        v = D.dftGS(ix, roots, x)
        assert abs(complex(v.item()) - want[ix]) <= tol * np.abs(want).max() * 4


def test_dft_rejects_aliasing_and_bad_index(gpu):
    x = dev(O.uniform(1, (8,)), torch.complex64, gpu)
    lib = afft._lib.lib()
    e = afft._lib.extents_arg([8])
    p = x.data_ptr()
    assert lib.afft_dft(0, 0, 1, e, p, p, None) == 1
    bad = afft._lib.extents_arg([8])
    assert lib.afft_dftGS(0, 1, e, bad, p, p, p, None) == 1  # index 8 of an extent 8
