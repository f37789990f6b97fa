"""Fused spectral convolution y = Inverse(fft2D(x) * H) and its VJP, against the oracle, against
the reference's own Reverse mode, and against central finite differences in FP64."""
import ctypes

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu


def dev(x, dtype, device):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dtype).to(device)


def host(t):
    return t.detach().cpu().numpy().astype(np.complex128)


def oracle_conv(x, H):
    return np.stack([O.fft2D(O.Mode.Inverse, O.fft2D(O.Mode.Forward, xb) * H) for xb in x])


def oracle_vjp(x, H, gy):
    M = H.size
    # (1/M) F^H (conj(H) F gy): the reference's Reverse mode divided by M, i.e. Inverse
    gx = np.stack([O.fft2D(O.Mode.Reverse, np.conj(H) * O.fft2D(O.Mode.Forward, g)) / M for g in gy])
    gH = sum(np.conj(O.fft2D(O.Mode.Forward, xb)) * O.fft2D(O.Mode.Forward, g) for xb, g in zip(x, gy)) / M
    return gx, gH


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("shape", [(3, 64, 96), (2, 512, 512), (5, 30, 1), (1, 1, 8), (4, 96, 64), (2, 256, 48)])
def test_spectral_conv_forward(gpu, dtype, shape):
    B, h, w = shape
    x = O.normal(1, shape)
    H = O.normal(2, (h, w))
    xd, Hd = dev(x, dtype, gpu), dev(H, dtype, gpu)
    y = afft.spectral_conv(xd, Hd)
    want = oracle_conv(host(xd), host(Hd))
    err = O.rel_l2(host(y), want)
    assert err <= tolerance(dtype == torch.complex128, h * w), err
    # the per-call composition on the GPU gives the same result
    plain = torch.stack([afft.fft2D(Mode.Inverse, afft.fft2D(Mode.Forward, xd[b]) * Hd) for b in range(B)])
    assert O.rel_l2(host(y), host(plain)) <= 2 * tolerance(dtype == torch.complex128, h * w)
    batched = afft.fft2D_many(Mode.Inverse, afft.fft2D_many(Mode.Forward, xd) * Hd)
    assert O.rel_l2(host(batched), want) <= tolerance(dtype == torch.complex128, h * w)


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
# h = 256 and 512 take the compile-time column transforms, the others the generic one
@pytest.mark.parametrize("shape", [(4, 64, 64), (7, 32, 48), (64, 128, 128), (3, 256, 40), (2, 512, 24)])
def test_spectral_conv_vjp_matches_oracle(gpu, dtype, shape):
    B, h, w = shape
    x, H, gy = O.normal(3, shape), O.normal(4, (h, w)), O.normal(5, shape)
    xd, Hd, gd = dev(x, dtype, gpu), dev(H, dtype, gpu), dev(gy, dtype, gpu)
    gx, gH = afft.spectral_conv_vjp(xd, Hd, gd)
    wx, wH = oracle_vjp(host(xd), host(Hd), host(gd))
    tol = tolerance(dtype == torch.complex128, h * w)
    assert O.rel_l2(host(gx), wx) <= tol
    assert O.rel_l2(host(gH), wH) <= tol * 4
    # only one of the two outputs requested
    gx2, none = afft.spectral_conv_vjp(xd, Hd, gd, want_H=False)
    assert none is None and O.rel_l2(host(gx2), wx) <= tol
    none, gH2 = afft.spectral_conv_vjp(xd, Hd, gd, want_x=False)
    assert none is None and O.rel_l2(host(gH2), wH) <= tol * 4


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
@pytest.mark.parametrize("shape", [(4, 64, 64), (64, 128, 128), (3, 14, 10), (2, 512, 24)])
def test_spectral_conv_saved_spectrum(gpu, dtype, shape):
    """The forward that stores F x and the VJP that reads it instead of transforming x (fused
    shapes, and h = 14, which is not 5-smooth: the unfused composition), against the recomputing
    forms and the oracle."""
    B, h, w = shape
    x, H, gy = O.normal(13, shape), O.normal(14, (h, w)), O.normal(15, shape)
    xd, Hd, gd = dev(x, dtype, gpu), dev(H, dtype, gpu), dev(gy, dtype, gpu)
    tol = tolerance(dtype == torch.complex128, h * w)
    y, spec = afft.spectral_conv(xd, Hd, save_spectrum=True)
    assert O.rel_l2(host(y), host(afft.spectral_conv(xd, Hd))) <= tol
    assert O.rel_l2(host(spec), np.stack([O.fft2D(O.Mode.Forward, xb) for xb in host(xd)])) <= tol
    keep = spec.clone()
    gx, gH = afft.spectral_conv_vjp(None, Hd, gd, spectrum=spec)
    assert torch.equal(spec, keep)  # read, not modified
    wx, wH = oracle_vjp(host(xd), host(Hd), host(gd))
    assert O.rel_l2(host(gx), wx) <= tol
    assert O.rel_l2(host(gH), wH) <= tol * 4
    rx, rH = afft.spectral_conv_vjp(xd, Hd, gd)
    assert O.rel_l2(host(gx), host(rx)) <= tol and O.rel_l2(host(gH), host(rH)) <= tol * 4
    gx2, none = afft.spectral_conv_vjp(None, Hd, gd, want_H=False, spectrum=spec)
    assert none is None and O.rel_l2(host(gx2), wx) <= tol
    none, gH2 = afft.spectral_conv_vjp(None, Hd, gd, want_x=False, spectrum=spec)
    assert none is None and O.rel_l2(host(gH2), wH) <= tol * 4


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
# w = 256 and 512 (and 1024 in FP32) take the residual-load row kernel, the others the stored
# difference; h = 14 is the unfused composition
@pytest.mark.parametrize("shape", [(4, 64, 512), (3, 256, 256), (2, 32, 1024), (5, 64, 96), (3, 14, 10)])
def test_spectral_conv_vjp_residual(gpu, dtype, shape):
    """The squared-error VJP (gy = alpha (y - t) formed on load) against the oracle's VJP of the
    same gy, from x and from the stored spectrum."""
    B, h, w = shape
    x, H, t = O.normal(23, shape), O.normal(24, (h, w)), O.normal(25, shape)
    xd, Hd, td = dev(x, dtype, gpu), dev(H, dtype, gpu), dev(t, dtype, gpu)
    tol = tolerance(dtype == torch.complex128, h * w)
    y, spec = afft.spectral_conv(xd, Hd, save_spectrum=True)
    for alpha in (2.0, -0.75):
        gy = alpha * (host(y) - host(td))
        wx, wH = oracle_vjp(host(xd), host(Hd), gy)
        for kw in ({"x": xd}, {"x": None, "spectrum": spec}):
            gx, gH = afft.spectral_conv_vjp_residual(kw["x"], Hd, y, td, alpha, spectrum=kw.get("spectrum"))
            assert O.rel_l2(host(gx), wx) <= tol, (alpha, kw.keys())
            assert O.rel_l2(host(gH), wH) <= tol * 4, (alpha, kw.keys())
    none, gH2 = afft.spectral_conv_vjp_residual(xd, Hd, y, td, 2.0, want_x=False)
    assert none is None
    assert O.rel_l2(host(gH2), oracle_vjp(host(xd), host(Hd), 2.0 * (host(y) - host(td)))[1]) <= tol * 4


def test_spectral_saved_spectrum_rejects_aliasing(gpu):
    lib = afft._lib.lib()
    x = torch.zeros(2, 16, 16, dtype=torch.complex64, device=gpu)
    H = torch.zeros(16, 16, dtype=torch.complex64, device=gpu)
    y = torch.empty_like(x)
    s = torch.cuda.current_stream().cuda_stream
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    assert lib.afft_spectral_conv_fwd_save(0, 2, 16, 16, p(x), p(H), p(y), p(y), ctypes.c_void_p(s)) == 1
    assert lib.afft_spectral_conv_fwd_save(0, 2, 16, 16, p(x), p(H), p(y), None, ctypes.c_void_p(s)) == 1
    assert lib.afft_spectral_conv_vjp_saved(0, 2, 16, 16, p(x), p(H), p(y), p(x), None, ctypes.c_void_p(s)) == 1


def test_spectral_conv_vjp_finite_differences_fp64(gpu):
    # loss = sum |y - t|^2 on 4 x 64 x 64, central differences in FP64 along random directions
    B, h, w = 4, 64, 64
    dt = torch.complex128
    x, H, t = dev(O.normal(6, (B, h, w)), dt, gpu), dev(O.normal(7, (h, w)), dt, gpu), dev(O.normal(8, (B, h, w)), dt, gpu)

    def loss(xx, HH):
        y = afft.spectral_conv(xx, HH)
        return float(torch.sum(torch.abs(y - t) ** 2))

    y = afft.spectral_conv(x, H)
    gy = 2 * (y - t)  # dL/d(conj y) times 2: the torch convention
    gx, gH = afft.spectral_conv_vjp(x, H, gy)
    eps = 1e-6
    for seed in range(3):
        dx = dev(O.normal(100 + seed, (B, h, w)), dt, gpu)
        dH = dev(O.normal(200 + seed, (h, w)), dt, gpu)
        fd_x = (loss(x + eps * dx, H) - loss(x - eps * dx, H)) / (2 * eps)
        fd_H = (loss(x, H + eps * dH) - loss(x, H - eps * dH)) / (2 * eps)
        an_x = float(torch.sum(torch.conj(gx) * dx).real)
        an_H = float(torch.sum(torch.conj(gH) * dH).real)
        assert abs(fd_x - an_x) <= 1e-6 * max(1.0, abs(an_x)), (fd_x, an_x)
        assert abs(fd_H - an_H) <= 1e-6 * max(1.0, abs(an_H)), (fd_H, an_H)


def test_spectral_configuration_size(gpu):
    # 64 x 512 x 512 Complex Float: fused forward and VJP against the per-call composition
    B, h, w = 64, 512, 512
    g = torch.Generator(device=gpu)
    g.manual_seed(5)
    x = torch.complex(torch.randn(B, h, w, generator=g, device=gpu), torch.randn(B, h, w, generator=g, device=gpu))
    H = torch.complex(torch.randn(h, w, generator=g, device=gpu), torch.randn(h, w, generator=g, device=gpu))
    gy = torch.complex(torch.randn(B, h, w, generator=g, device=gpu), torch.randn(B, h, w, generator=g, device=gpu))
    y = afft.spectral_conv(x, H)
    plain = torch.stack([afft.fft2D(Mode.Inverse, afft.fft2D(Mode.Forward, x[b]) * H) for b in range(B)])
    tol = tolerance(False, h * w)
    assert O.rel_l2(host(y), host(plain)) <= tol
    gx, gH = afft.spectral_conv_vjp(x, H, gy)
    Fx = torch.stack([afft.fft2D(Mode.Forward, x[b]) for b in range(B)])
    Fg = torch.stack([afft.fft2D(Mode.Forward, gy[b]) for b in range(B)])
    px = torch.stack([afft.fft2D(Mode.Inverse, torch.conj(H) * Fg[b]) for b in range(B)])
    pH = torch.sum(torch.conj(Fx) * Fg, dim=0) / (h * w)
    assert O.rel_l2(host(gx), host(px)) <= tol
    assert O.rel_l2(host(gH), host(pH)) <= tol
    # two images against the FP64 oracle
    xs = host(x[:2])
    want = oracle_conv(xs, host(H))
    assert O.rel_l2(host(y[:2]), want) <= tol
