"""The benchmark's configurations at their full sizes against the FP64 oracle: every element of
every output, on the benchmark's own seeded inputs (SplitMix64 seeds 2-7), in every mode the
configurations name.  The composed multi-pass plans the benchmark times are checked here as they
run there, not through smaller shapes."""
import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu

OMODE = {Mode.Forward: O.Mode.Forward, Mode.Reverse: O.Mode.Reverse, Mode.Inverse: O.Mode.Inverse}


def on_gpu(x, gpu):
    return torch.from_numpy(np.ascontiguousarray(x)).to(gpu)


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Reverse, Mode.Inverse])
def test_fft_16384x4096_every_row(gpu, mode):
    # batched fft along the innermost axis of Array DIM2 (Complex Float) 16384 x 4096, uniform, seed 2
    xh = O.uniform(2, (16384, 4096), np.complex64)
    y = afft.fft(mode, on_gpu(xh, gpu)).cpu().numpy()
    want = O.fft(OMODE[mode], xh)
    tol = tolerance(False, 4096)
    assert O.rel_l2(y, want) <= tol
    # and row by row: no single transform beyond the bound
    num = np.linalg.norm(y - want, axis=1)
    den = np.linalg.norm(want, axis=1)
    assert float(np.max(num / den)) <= tol


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Reverse, Mode.Inverse])
def test_fft_1x4194304(gpu, mode):
    xh = O.uniform(2, (1, 1 << 22), np.complex64)
    y = afft.fft(mode, on_gpu(xh, gpu)).cpu().numpy()
    assert O.rel_l2(y, O.fft(OMODE[mode], xh)) <= tolerance(False, 1 << 22)


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Inverse])
def test_fft2D_4096x4096(gpu, mode):
    xh = O.normal(3, (4096, 4096), np.complex64)
    y = afft.fft2D(mode, on_gpu(xh, gpu)).cpu().numpy()
    assert O.rel_l2(y, O.fft2D(OMODE[mode], xh)) <= tolerance(False, 4096 * 4096)


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Inverse])
def test_fft3D_256_cubed_complex_double(gpu, mode):
    xh = O.normal(3, (256, 256, 256))
    y = afft.fft3D(mode, on_gpu(xh, gpu)).cpu().numpy()
    assert O.rel_l2(y, O.fft3D(OMODE[mode], xh)) <= tolerance(True, 256 ** 3)


def test_fft2D_many_64x512x512(gpu):
    xh = O.normal(5, (64, 512, 512), np.complex64)
    y = afft.fft2D_many(Mode.Forward, on_gpu(xh, gpu)).cpu().numpy()
    assert O.rel_l2(y, O.fft_axes(xh, (1, 2), -1)) <= tolerance(False, 512 * 512)


def test_spectral_configuration_all_64_images(gpu):
    # y = Inverse(fft2D(x) * H), loss = sum |y - t|^2: x 64 x 512 x 512, H 512 x 512, t, all
    # normal (seeds 5, 6, 7); the gradients from gy = 2 (y - t), every VJP form
    B, h, w = 64, 512, 512
    xh = O.normal(5, (B, h, w), np.complex64)
    Hh = O.normal(6, (h, w), np.complex64)
    th = O.normal(7, (B, h, w), np.complex64)
    x, H, t = on_gpu(xh, gpu), on_gpu(Hh, gpu), on_gpu(th, gpu)
    tol = tolerance(False, h * w)
    y, spec = afft.spectral_conv(x, H, save_spectrum=True)
    gy = 2 * (y - t)
    gx, gH = afft.spectral_conv_vjp(x, H, gy)
    sx, sH = afft.spectral_conv_vjp(None, H, gy, spectrum=spec)
    # the squared-error form (gy = 2 (y - t) formed on load), from x and from the spectrum
    rx, rH = afft.spectral_conv_vjp_residual(x, H, y, t, 2.0)
    qx, qH = afft.spectral_conv_vjp_residual(None, H, y, t, 2.0, spectrum=spec)
    Fx = np.stack([O.fft2D(O.Mode.Forward, xb) for xb in xh])
    assert O.rel_l2(spec.cpu().numpy(), Fx) <= tol
    wy = np.stack([O.fft2D(O.Mode.Inverse, f * Hh) for f in Fx])
    assert O.rel_l2(y.cpu().numpy(), wy) <= tol
    g = gy.cpu().numpy().astype(np.complex128)
    Fg = np.stack([O.fft2D(O.Mode.Forward, gb) for gb in g])
    wgx = np.stack([O.fft2D(O.Mode.Reverse, np.conj(Hh) * f) / (h * w) for f in Fg])
    wgH = np.sum(np.conj(Fx) * Fg, axis=0) / (h * w)
    for got_x, got_H in ((gx, gH), (sx, sH), (rx, rH), (qx, qH)):
        assert O.rel_l2(got_x.cpu().numpy(), wgx) <= tol
        assert O.rel_l2(got_H.cpu().numpy(), wgH) <= tol
    # image by image: every image of gx within the bound
    gxh = gx.cpu().numpy()
    per = np.linalg.norm((gxh - wgx).reshape(B, -1), axis=1) / np.linalg.norm(wgx.reshape(B, -1), axis=1)
    assert float(np.max(per)) <= tol
