"""afft.centre -- host-side mirror of ``Data.Array.Accelerate.Math.DFT.Centre`` over libafft.so.

The nine functions of DFT/Centre.hs:21-26 on rank-1/2/3 device arrays (complex64 / complex128),
each returning a fresh array:

* ``centre1D/2D/3D``  x * (-1)^(sum of the multi-index)                  (Centre.hs:36-65)
* ``shift1D/2D/3D``   y[i] = x[(i + ceil(n/2)) mod n] along every axis    (Centre.hs:70-82, 104-123)
* ``ishift1D/2D/3D``  y[i] = x[(i + floor(n/2)) mod n] along every axis   (Centre.hs:84-101, 125-164)

and ``fft_centred(mode, x, pre, post)`` = post(fftND(mode, pre(x))) with the centring folded into
the transform (afft_fft_centred): ``fft_centred(Forward, x, pre="centre")`` is
``fft2D Forward (centre2D x)`` in one call with no extra pass.  The reference's identity
``fft(centre(X)) = shift(fft(X))`` (Centre.hs:17-19) holds for even extents.
"""
import torch

from . import _lib as L
from . import _device_tensor, _dtype_code, _mode, _ptr, _stream

_WHICH = {None: L.REORDER_NONE, "none": L.REORDER_NONE, "centre": L.CENTRE, "shift": L.SHIFT,
          "ishift": L.ISHIFT}


def _reorder(which, arr, rank):
    x = _device_tensor(arr)
    if x.dim() != rank:
        raise ValueError(f"expected an Array DIM{rank}")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_reorder(which, _dtype_code(x), x.dim(), L.extents_arg(x.shape), _ptr(x),
                                     _ptr(out), _stream(x.device)), "afft_reorder")
    return out


def centre1D(x): return _reorder(L.CENTRE, x, 1)
def centre2D(x): return _reorder(L.CENTRE, x, 2)
def centre3D(x): return _reorder(L.CENTRE, x, 3)
def shift1D(x): return _reorder(L.SHIFT, x, 1)
def shift2D(x): return _reorder(L.SHIFT, x, 2)
def shift3D(x): return _reorder(L.SHIFT, x, 3)
def ishift1D(x): return _reorder(L.ISHIFT, x, 1)
def ishift2D(x): return _reorder(L.ISHIFT, x, 2)
def ishift3D(x): return _reorder(L.ISHIFT, x, 3)


def fft_centred(mode, arr, pre=None, post=None, out=None):
    """post(fftND(mode, pre(arr))) over all axes of a rank-1/2/3 array; pre / post are None,
    "centre", "shift" or "ishift".  ``out`` may be ``arr`` itself (in place)."""
    x = _device_tensor(arr)
    if not 1 <= x.dim() <= 3:
        raise ValueError("fft_centred takes an Array DIM1, DIM2 or DIM3")
    if out is None:
        out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_fft_centred(_mode(mode), _dtype_code(x), x.dim(), L.extents_arg(x.shape),
                                         _WHICH[pre], _WHICH[post], _ptr(x), _ptr(out),
                                         _stream(x.device)), "afft_fft_centred")
    return out
