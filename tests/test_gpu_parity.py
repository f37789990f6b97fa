"""GPU parity through the C ABI: every result against the FP64 oracle on the same seeded inputs,
at relative-L2 <= 1e-13 (Complex Double) or 1e-6*log2(N) (Complex Float)."""
import ctypes
import math
import os
import threading

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dft_mpmath.npz")
DTYPES = [torch.complex64, torch.complex128]
OMODE = {Mode.Forward: O.Mode.Forward, Mode.Reverse: O.Mode.Reverse, Mode.Inverse: O.Mode.Inverse}


def dev(x, dtype, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).to(device)


def host(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def check(got, want, dtype, n, what=""):
    err = O.rel_l2(host(got), want)
    tol = tolerance(dtype == torch.complex128, n)
    assert err <= tol, f"{what}: rel L2 {err:.3e} > {tol:.3e}"
    return err


@pytest.fixture(scope="module")
def golden():
    with np.load(GOLDEN, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fft1D_complex_double_1024_seed0(gpu, golden):
    # fft1D Forward then Inverse on Array DIM1 (Complex Double), N=1024, seed 0
    x = golden["x_1024"]
    np.testing.assert_array_equal(x, O.uniform(0, (1024,)))
    xd = dev(x, torch.complex128, gpu)
    X = afft.fft1D(Mode.Forward, xd)
    check(X, golden["fwd_1024"], torch.complex128, 1024, "forward vs mpmath")
    check(X, O.fft1D(O.Mode.Forward, x), torch.complex128, 1024, "forward vs oracle")
    back = afft.fft1D(Mode.Inverse, X)
    check(back, x, torch.complex128, 1024, "round trip")
    np.testing.assert_array_equal(host(xd), x)  # input preserved


@pytest.mark.parametrize("dtype", DTYPES)
def test_golden_vectors(gpu, golden, dtype):
    ns = sorted(int(k[2:]) for k in golden if k.startswith("x_"))
    for n in ns:
        x = golden[f"x_{n}"]
        xd = dev(x, dtype, gpu)
        want_x = host(xd)  # FP32 inputs are the rounded values
        check(afft.fft1D(Mode.Forward, xd), O.fft1D(O.Mode.Forward, want_x), dtype, n, f"F n={n}")
        if dtype == torch.complex128:
            check(afft.fft1D(Mode.Forward, xd), golden[f"fwd_{n}"], dtype, n, f"F mpmath n={n}")
            check(afft.fft1D(Mode.Reverse, xd), golden[f"rev_{n}"], dtype, n, f"R mpmath n={n}")
        check(afft.fft1D(Mode.Inverse, xd), O.fft1D(O.Mode.Inverse, want_x), dtype, n, f"I n={n}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_every_length_1_to_1024(gpu, dtype):
    # the DIM1 generator range of test/Test/Base.hs:44-45, every length
    for n in range(1, 1025):
        x = O.uniform(n, (n,), np.complex128)
        xd = dev(x, dtype, gpu)
        xr = host(xd)
        mode = (Mode.Forward, Mode.Reverse, Mode.Inverse)[n % 3]
        check(afft.fft1D(mode, xd), O.fft1D(OMODE[mode], xr), dtype, n, f"n={n} {mode}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(48, 128), (17, 100), (1, 1), (5, 1), (3, 97), (16, 32, 64),
                                   (7, 9, 60), (2, 3, 4, 5, 6), (1000, 64), (4, 4096)])
def test_fft_innermost_batched(gpu, dtype, shape):
    x = O.uniform(sum(shape), shape)
    xd = dev(x, dtype, gpu)
    xr = host(xd)
    for mode in Mode:
        check(afft.fft(mode, xd), O.fft(OMODE[mode], xr), dtype, shape[-1], f"{shape} {mode}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(48, 128), (1, 1), (1, 64), (64, 1), (13, 7), (30, 24), (256, 256),
                                   (97, 64), (512, 384)])
def test_fft2D(gpu, dtype, shape):
    x = O.normal(7 + sum(shape), shape)
    xd = dev(x, dtype, gpu)
    xr = host(xd)
    n = shape[0] * shape[1]
    for mode in Mode:
        check(afft.fft2D(mode, xd), O.fft2D(OMODE[mode], xr), dtype, n, f"{shape} {mode}")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(4096, 64), (2048, 32), (3072, 40), (4096, 16), (1024, 128),
                                   (6000, 20), (4096, 4, 32)])
def test_long_strided_axes(gpu, dtype, shape):
    """Columns too long for a line-wide single-workgroup tile: split four-step into passes of
    register-resident (16/32/64) or LDS transforms, each with whole-line row segments."""
    x = O.normal(5 + sum(shape), shape)
    xd = dev(x, dtype, gpu)
    xr = host(xd)
    n = int(np.prod(shape))
    fn, ofn = (afft.fft2D, O.fft2D) if len(shape) == 2 else (afft.fft3D, O.fft3D)
    for mode in Mode:
        check(fn(mode, xd), ofn(OMODE[mode], xr), dtype, n, f"{shape} {mode}")
    # in place on the output buffer (rows, then columns through the work buffer)
    y = xd.clone()
    p = afft.Plan(shape, dtype, "all")
    p.execute(y, out=y)
    check(y, O.fft_axes(xr, tuple(range(len(shape))), -1), dtype, n, f"{shape} in place")


def test_two_split_axes(gpu):
    """Complex Double 2048 x 16384: the rows split six-step (32 x 512) and the 2048-long columns
    split four-step (64-byte tiles otherwise), so the plan's first pass writes the output, the
    row transpose moves the data to the work buffer and the column pass B brings it back."""
    shape = (2048, 16384)
    x = O.normal(21, shape)
    xd = dev(x, torch.complex128, gpu)
    n = shape[0] * shape[1]
    want = O.fft2D(O.Mode.Forward, x)
    check(afft.fft2D(Mode.Forward, xd), want, torch.complex128, n, "two splits")
    p = afft.Plan(shape, torch.complex128, "all")
    assert p.info()["launches"] == 5  # A, rows, transpose; column A, column B
    y = xd.clone()
    p.execute(y, out=y)
    check(y, want, torch.complex128, n, "two splits in place")


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("shape", [(16, 32, 64), (1, 1, 1), (3, 5, 7), (8, 1, 8), (11, 12, 10),
                                   (32, 32, 32), (64, 16, 48), (256, 4, 8), (4, 256, 8)])
def test_fft3D(gpu, dtype, shape):
    x = O.normal(11 + sum(shape), shape)
    xd = dev(x, dtype, gpu)
    xr = host(xd)
    n = shape[0] * shape[1] * shape[2]
    for mode in Mode:
        check(afft.fft3D(mode, xd), O.fft3D(OMODE[mode], xr), dtype, n, f"{shape} {mode}")


@pytest.mark.parametrize("dtype,n", [(torch.complex64, 32768), (torch.complex64, 1 << 18),
                                     (torch.complex64, 3 * 8192 * 4), (torch.complex128, 16384),
                                     (torch.complex128, 40960), (torch.complex64, 1 << 22),
                                     (torch.complex64, 1 << 20), (torch.complex128, 1 << 20),
                                     (torch.complex64, 5 * 64 * 64 * 512)])
def test_long_rows_four_step(gpu, dtype, n):
    x = O.uniform(n, (2, n))
    xd = dev(x, dtype, gpu)
    xr = host(xd)
    for mode in Mode:
        check(afft.fft(mode, xd), O.fft(OMODE[mode], xr), dtype, n, f"n={n} {mode}")


@pytest.mark.parametrize("shape", [(1, 1 << 23), (1, 1 << 22), (1, 1 << 21), (1, 3 << 20), (1, 5 << 19),
                                   (3, 1 << 20)])
def test_long_rows_split_six_step(gpu, shape):
    """Rows whose six-step row length has no compiled row kernel: a register column pass (64, 32
    or 16 points, by the thread count), a strided LDS pass, and transposed-store rows of 512, 768
    or 640 read through a row table -- three launches."""
    n = shape[-1]
    x = O.uniform(11, shape)
    xd = dev(x, torch.complex64, gpu)
    xr = host(xd)
    p = afft.Plan(shape, torch.complex64)
    assert p.info()["launches"] == 3
    for mode in Mode:
        check(afft.fft(mode, xd), O.fft(OMODE[mode], xr), torch.complex64, n, f"{shape} {mode}")
    y = xd.clone()
    p.execute(y, direction=+1, scale=1.0 / n, out=y)  # Inverse, in place
    check(y, O.fft(O.Mode.Inverse, xr), torch.complex64, n, f"{shape} in place")
    p.close()


def _smooth_lengths(lo, hi):
    out = []
    for n in range(lo, hi + 1):
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1:
            out.append(n)
    return out


@pytest.mark.parametrize("dtype,stride", [(torch.complex64, 8), (torch.complex128, 24)])
def test_long_row_lengths_sweep(gpu, dtype, stride):
    """Every stride-th 5-smooth length in (16384, 2^21], odd ones included: single-workgroup,
    six-step, split six-step and balanced four-step plans, whichever each length selects."""
    lengths = _smooth_lengths(16385, 1 << 21)[::stride]
    for n in lengths:
        x = O.uniform(n % 1000, (1, n))
        xd = dev(x, dtype, gpu)
        check(afft.fft(Mode.Forward, xd), O.fft(O.Mode.Forward, host(xd)), dtype, n, f"n={n}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_bluestein_primes_and_ragged_lengths(gpu, dtype):
    # 4099 (Complex Double) and 10007 pad beyond one workgroup: the multi-kernel path
    for n in (257, 1009, 4099, 3 * 7 * 11, 2 * 1021, 10007):
        x = O.uniform(n, (3, n))
        xd = dev(x, dtype, gpu)
        xr = host(xd)
        for mode in Mode:
            check(afft.fft(mode, xd), O.fft(OMODE[mode], xr), dtype, n, f"n={n} {mode}")


@pytest.mark.parametrize("dtype", DTYPES)
def test_out_of_place_preserves_input_and_in_place_works(gpu, dtype):
    lib = afft._lib.lib()
    code = 0 if dtype == torch.complex64 else 1
    for shape in [(8, 4096), (1, 1 << 16), (64, 200)]:
        x = O.uniform(5, shape)
        xd = dev(x, dtype, gpu)
        keep = xd.clone()
        y = afft.fft(Mode.Forward, xd)
        assert torch.equal(xd, keep)
        ext = afft._lib.extents_arg(shape)
        z = xd.clone()
        st = lib.afft_fft(0, code, len(shape), ext, ctypes.c_void_p(z.data_ptr()), ctypes.c_void_p(z.data_ptr()),
                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st == 0
        assert O.rel_l2(host(z), host(y)) <= tolerance(dtype == torch.complex128, shape[-1])


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("op,shape", [("fft", (6, 100)), ("fft1D", (1024,)), ("fft2D", (24, 40)), ("fft3D", (4, 6, 10))])
def test_vjp_is_the_adjoint(gpu, dtype, op, shape):
    T = getattr(afft, op)
    x = dev(O.uniform(1, shape), dtype, gpu)
    y = dev(O.uniform(2, shape), dtype, gpu)
    for mode in Mode:
        lhs = torch.vdot(T(mode, x).flatten(), y.flatten())        # <F x, y>
        rhs = torch.vdot(x.flatten(), afft.vjp(op, mode, y).flatten())  # <x, F^H y>
        rel = abs(complex(lhs) - complex(rhs)) / max(1e-30, abs(complex(lhs)))
        assert rel < (1e-12 if dtype == torch.complex128 else 1e-5), (op, mode, rel)


def test_explicit_plan_api(gpu):
    x = dev(O.uniform(9, (32, 4096)), torch.complex64, gpu)
    p = afft.Plan((32, 4096), torch.complex64, "innermost")
    info = p.info()
    assert info["launches"] == 1 and info["device_bytes"] > 0
    y = p.execute(x, -1, 1.0)
    check(y, O.fft(O.Mode.Forward, host(x)), torch.complex64, 4096)
    z = p.execute(y, +1, 1.0 / 4096)
    check(z, host(x), torch.complex64, 4096)
    p.close()
    # rows, then the 4096-long columns as two register-resident passes of 64 (four-step)
    q = afft.Plan((4096, 4096), torch.complex64, "all")
    assert q.info()["launches"] == 3
    q.close()


def test_concurrent_host_threads_on_separate_streams(gpu):
    shapes = [(64, 1024), (32, 2048), (16, 4096), (64, 1000)]
    xs = [dev(O.uniform(i, s), torch.complex64, gpu) for i, s in enumerate(shapes)]
    want = [O.fft(O.Mode.Forward, host(x)) for x in xs]
    errors = []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(20):
                    y = afft.fft(Mode.Forward, xs[i])
                s.synchronize()
            if O.rel_l2(host(y), want[i]) > 2e-5:
                errors.append(i)
        except Exception as e:  # pragma: no cover
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(shapes))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors


def test_errors_are_statuses(gpu):
    with pytest.raises(TypeError):
        afft.fft(Mode.Forward, torch.zeros(8, device=gpu))
    with pytest.raises(ValueError):
        afft.fft2D(Mode.Forward, torch.zeros(8, dtype=torch.complex64, device=gpu))
    # a 5-smooth length beyond two single-workgroup factors (16384^2 for Complex Float) is
    # refused at planning time, before any allocation or launch
    lib = afft._lib.lib()
    buf = torch.zeros(4, dtype=torch.complex64, device=gpu)
    ext = afft._lib.extents_arg([3 << 29])
    st = lib.afft_fft(0, 0, 1, ext, ctypes.c_void_p(buf.data_ptr()), ctypes.c_void_p(buf.data_ptr()),
                      ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st == 4
