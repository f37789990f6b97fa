"""afft -- host-side mirror of ``Data.Array.Accelerate.Math.FFT`` over the gfx950 FFT engine.

The names, argument meaning and errors follow the reference front-end
(src/Data/Array/Accelerate/Math/FFT.hs:25-35):

* ``Mode.Forward`` / ``Mode.Reverse`` / ``Mode.Inverse``  (FFT/Mode.hs:15-19)
* ``fft(mode, arr)``    DFT along the innermost axis of a rank >= 1 array, outer axes a batch
                        (FFT.hs:63-84); Inverse divides by the innermost extent.
* ``fft1D(mode, vec)``  (FFT.hs:92-111); Inverse divides by the length.
* ``fft2D(mode, mat)``  (FFT.hs:119-142); Inverse divides by the total size.
* ``fft3D(mode, arr)``  (FFT.hs:150-173); Inverse divides by the total size.

Arrays are torch tensors of dtype ``complex64`` (Complex Float) or ``complex128`` (Complex
Double) resident on a ROCm device; their storage is the interleaved Vec2 layout of
FFT/Type.hs:32.  Every call returns a fresh output tensor and leaves the input untouched, as
the plugins do (LLVM/PTX.hs:92, LLVM/Native/Base.hs:47).  The work is enqueued on the current
torch stream of the input's device, asynchronously.

torch is used here only for device memory and streams; every transform runs in libafft.so
(include/afft.h) and a missing library raises ``AfftError`` -- there is no fallback path.
"""
import ctypes
import enum

import torch

from . import _lib as L
from ._lib import AfftError

__all__ = ["Mode", "fft", "fft1D", "fft2D", "fft3D", "fft2D_many", "vjp", "Plan", "AfftError",
           "RaggedPlan", "spectral_conv", "spectral_conv_vjp", "describe", "centre", "dft"]


class Mode(enum.Enum):
    """Transform direction (FFT/Mode.hs:15-19)."""
    Forward = L.FORWARD   # exp(-2 pi i nk/N)
    Reverse = L.REVERSE   # exp(+2 pi i nk/N), un-normalised
    Inverse = L.INVERSE   # exp(+2 pi i nk/N) / N


def _dtype_code(t):
    if t.dtype == torch.complex64:
        return L.C64
    if t.dtype == torch.complex128:
        return L.C128
    # Numeric e is Float or Double only (FFT/Type.hs:28-39)
    raise TypeError(f"expected complex64 or complex128, got {t.dtype}")


def _device_tensor(arr):
    if not isinstance(arr, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    if arr.device.type != "cuda":
        raise ValueError("expected a tensor on a ROCm device")
    _dtype_code(arr)
    # a conjugated or negated view keeps the original storage and only a flag: materialise it,
    # or the library would read the unconjugated values behind the view
    return arr.resolve_conj().resolve_neg().contiguous()


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _mode(mode):
    return Mode(mode).value if not isinstance(mode, Mode) else mode.value


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def fft(mode, arr):
    """FFT.hs:63-84: DFT along the innermost axis; the outer axes form a batch."""
    x = _device_tensor(arr)
    if x.dim() < 1:
        raise ValueError("fft needs an array of rank >= 1 (sh :. Int)")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft(_mode(mode), _dtype_code(x), x.dim(), L.extents_arg(x.shape),
                                 _ptr(x), _ptr(out), _stream(x.device)), "afft_fft")
    return out


def fft1D(mode, arr):
    """FFT.hs:92-111."""
    x = _device_tensor(arr)
    if x.dim() != 1:
        raise ValueError("fft1D takes an Array DIM1")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft1D(_mode(mode), _dtype_code(x), x.shape[0], _ptr(x), _ptr(out),
                                   _stream(x.device)), "afft_fft1D")
    return out


def fft2D(mode, arr):
    """FFT.hs:119-142."""
    x = _device_tensor(arr)
    if x.dim() != 2:
        raise ValueError("fft2D takes an Array DIM2")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft2D(_mode(mode), _dtype_code(x), x.shape[0], x.shape[1], _ptr(x),
                                   _ptr(out), _stream(x.device)), "afft_fft2D")
    return out


def fft3D(mode, arr):
    """FFT.hs:150-173."""
    x = _device_tensor(arr)
    if x.dim() != 3:
        raise ValueError("fft3D takes an Array DIM3")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft3D(_mode(mode), _dtype_code(x), x.shape[0], x.shape[1],
                                   x.shape[2], _ptr(x), _ptr(out), _stream(x.device)), "afft_fft3D")
    return out


def fft2D_many(mode, arr):
    """fft2D of every h x w matrix of a (batch, h, w) array in one call (extension: the
    reference has no batched 2D entry point, FFT.hs:119-122); Inverse divides by h*w."""
    x = _device_tensor(arr)
    if x.dim() != 3:
        raise ValueError("fft2D_many takes a (batch, h, w) array")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft2D_many(_mode(mode), _dtype_code(x), x.shape[0], x.shape[1],
                                        x.shape[2], _ptr(x), _ptr(out), _stream(x.device)),
                "afft_fft2D_many")
    return out


_OPS = {"fft": 0, "fft1D": 1, "fft2D": 2, "fft3D": 3}


def vjp(op, mode, gy):
    """Cotangent of the input of ``op(mode, x)`` given the output cotangent ``gy``
    (conjugate-Wirtinger convention, as torch autograd): op^H(gy)."""
    g = _device_tensor(gy)
    code = _OPS[op] if isinstance(op, str) else int(op)
    out = torch.empty_like(g)
    with torch.cuda.device(g.device):
        L.check(L.lib().afft_fft_vjp(code, _mode(mode), _dtype_code(g), g.dim(),
                                     L.extents_arg(g.shape), _ptr(g), _ptr(out),
                                     _stream(g.device)), "afft_fft_vjp")
    return out


def describe(n, dtype=torch.complex64):
    """Host-only: (radices, elements per thread, Bluestein length or 0) of the
    single-workgroup layout for length n; raises AfftError beyond the LDS limit."""
    code = L.C64 if dtype == torch.complex64 else L.C128
    rad = (ctypes.c_int * 32)()
    npass, E = ctypes.c_int(), ctypes.c_int()
    blue = ctypes.c_int64()
    L.check(L.lib().afft_describe(int(n), code, rad, 32, ctypes.byref(npass), ctypes.byref(E),
                                  ctypes.byref(blue)), "afft_describe")
    return list(rad[: npass.value]), E.value, blue.value


class Plan:
    """An explicit plan (afft_plan_create); ``axes`` is "innermost" (fft) or "all"."""

    def __init__(self, shape, dtype=torch.complex64, axes="innermost", device=None):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        code = L.C64 if dtype == torch.complex64 else L.C128
        ax = L.AXES_INNERMOST if axes == "innermost" else L.AXES_ALL
        self._h = ctypes.c_void_p()
        L.check(L.lib().afft_plan_create(ctypes.byref(self._h), code, len(self.shape),
                                         L.extents_arg(self.shape), ax, self.device.index),
                "afft_plan_create")

    def execute(self, x, direction=-1, scale=1.0, out=None):
        x = _device_tensor(x)
        if tuple(x.shape) != self.shape or x.dtype != self.dtype:
            raise ValueError("tensor does not match the plan")
        if out is None:
            out = torch.empty_like(x)
        with torch.cuda.device(x.device):
            L.check(L.lib().afft_exec(self._h, int(direction), _ptr(x), _ptr(out), float(scale),
                                      _stream(x.device)), "afft_exec")
        return out

    def info(self):
        n = ctypes.c_int()
        b = ctypes.c_int64()
        L.check(L.lib().afft_plan_info(self._h, ctypes.byref(n), ctypes.byref(b)), "afft_plan_info")
        return {"launches": n.value, "device_bytes": b.value}

    def close(self):
        if self._h:
            L.lib().afft_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RaggedPlan:
    """A ragged batch of 1D transforms of mixed lengths stored back to back, run by one
    persistent kernel (afft_ragged_plan_create / afft_ragged_exec)."""

    def __init__(self, lengths, dtype=torch.complex64, device=None):
        self.lengths = [int(n) for n in lengths]
        self.dtype = dtype
        self.device = torch.device("cuda", torch.cuda.current_device() if device is None else device)
        code = L.C64 if dtype == torch.complex64 else L.C128
        arr = (ctypes.c_int64 * max(1, len(self.lengths)))(*self.lengths)
        self._h = ctypes.c_void_p()
        L.check(L.lib().afft_ragged_plan_create(ctypes.byref(self._h), code, len(self.lengths), arr,
                                                self.device.index), "afft_ragged_plan_create")
        self.total = sum(self.lengths)

    def execute(self, mode, x, out=None, graph=False):
        x = _device_tensor(x)
        if x.numel() != self.total or x.dtype != self.dtype:
            raise ValueError("tensor does not match the ragged plan")
        if out is None:
            out = torch.empty_like(x)
        fn = L.lib().afft_ragged_exec_graph if graph else L.lib().afft_ragged_exec
        with torch.cuda.device(x.device):
            L.check(fn(self._h, _mode(mode), _ptr(x), _ptr(out), _stream(x.device)), "afft_ragged_exec")
        return out

    def close(self):
        if self._h:
            L.lib().afft_ragged_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def spectral_conv(x, H, save_spectrum=False):
    """y_b = Inverse(fft2D(x_b) * H) for a batch x of shape (B, h, w) and filter H (h, w).
    With save_spectrum, returns (y, F x): the spectrum the VJP can take instead of x."""
    x = _device_tensor(x)
    H = _device_tensor(H)
    if x.dim() != 3 or tuple(H.shape) != tuple(x.shape[1:]) or H.dtype != x.dtype:
        raise ValueError("expected x (B, h, w) and H (h, w) of the same dtype")
    y = torch.empty_like(x)
    with torch.cuda.device(x.device):
        if save_spectrum:
            spec = torch.empty_like(x)
            L.check(L.lib().afft_spectral_conv_fwd_save(_dtype_code(x), x.shape[0], x.shape[1], x.shape[2],
                                                        _ptr(x), _ptr(H), _ptr(y), _ptr(spec),
                                                        _stream(x.device)), "afft_spectral_conv_fwd_save")
            return y, spec
        L.check(L.lib().afft_spectral_conv_fwd(_dtype_code(x), x.shape[0], x.shape[1], x.shape[2],
                                               _ptr(x), _ptr(H), _ptr(y), _stream(x.device)),
                "afft_spectral_conv_fwd")
    return y


def spectral_conv_vjp(x, H, gy, want_x=True, want_H=True, spectrum=None):
    """(gx, gH) for y = spectral_conv(x, H) given dL/dy = gy (conjugate-Wirtinger).  F x is
    recomputed from x, or taken from `spectrum` (as spectral_conv(..., save_spectrum=True) returns
    it; x may then be None)."""
    H = _device_tensor(H)
    gy = _device_tensor(gy)
    if spectrum is not None:
        src = _device_tensor(spectrum)
        fn, name = L.lib().afft_spectral_conv_vjp_saved, "afft_spectral_conv_vjp_saved"
    else:
        src = _device_tensor(x)
        fn, name = L.lib().afft_spectral_conv_vjp, "afft_spectral_conv_vjp"
    if src.shape != gy.shape or src.dtype != gy.dtype:
        raise ValueError("x (or the spectrum) and gy must have the same shape and dtype")
    gx = torch.empty_like(gy) if want_x else None
    gH = torch.empty_like(H) if want_H else None
    with torch.cuda.device(gy.device):
        L.check(fn(_dtype_code(gy), gy.shape[0], gy.shape[1], gy.shape[2], _ptr(src), _ptr(H), _ptr(gy),
                   _ptr(gx) if gx is not None else None, _ptr(gH) if gH is not None else None,
                   _stream(gy.device)), name)
    return gx, gH


def spectral_conv_vjp_residual(x, H, y, t, alpha=2.0, want_x=True, want_H=True, spectrum=None):
    """(gx, gH) for the squared-error loss L = (alpha/2) sum |y - t|^2 of y = spectral_conv(x, H):
    spectral_conv_vjp with gy = alpha (y - t), formed as the first row pass loads y and t (gy is
    never stored).  F x is recomputed from x, or taken from `spectrum` (x may then be None)."""
    H = _device_tensor(H)
    y = _device_tensor(y)
    t = _device_tensor(t)
    src = _device_tensor(spectrum if spectrum is not None else x)
    if src.shape != y.shape or t.shape != y.shape or src.dtype != y.dtype or t.dtype != y.dtype:
        raise ValueError("x (or the spectrum), y and t must have the same shape and dtype")
    gx = torch.empty_like(y) if want_x else None
    gH = torch.empty_like(H) if want_H else None
    with torch.cuda.device(y.device):
        L.check(L.lib().afft_spectral_conv_vjp_residual(
            _dtype_code(y), y.shape[0], y.shape[1], y.shape[2], _ptr(src), 1 if spectrum is not None else 0,
            _ptr(H), _ptr(y), _ptr(t), float(alpha), _ptr(gx) if gx is not None else None,
            _ptr(gH) if gH is not None else None, _stream(y.device)), "afft_spectral_conv_vjp_residual")
    return gx, gH


from . import centre  # noqa: E402  (DFT/Centre.hs mirror; uses the helpers above)
from . import dft  # noqa: E402  (Math/DFT.hs mirror)
