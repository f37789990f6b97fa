"""afft.dft -- host-side mirror of ``Data.Array.Accelerate.Math.DFT`` (O(N^2)) over libafft.so.

* ``dft(x)``, ``idft(x)``   DFT.hs:42-60: the roots of unity of DFT/Roots.hs, idft divides by n
* ``dftG(roots, x)``        DFT.hs:63-92 with caller-supplied roots of x's shape
* ``dftGS(ix, roots, x)``   DFT.hs:95-111: the single value at index ``ix`` (a 0-d tensor)
* ``rootsOfUnity(shape)``, ``inverseRootsOfUnity(shape)``  Roots.hs:26-51

As in the reference, the sum of dftG runs over every element of the array (``arr !! i`` over
``shapeSize``, DFT.hs:79-92): for a vector it is the DFT; for a higher rank every row receives the
same values.  The transforms run on the device (afft_dft / afft_dftG / afft_dftGS).
"""
import math

import numpy as np
import torch

from . import _lib as L
from . import _device_tensor, _dtype_code, _ptr, _stream


def _run_dft(mode, arr):
    x = _device_tensor(arr)
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_dft(mode, _dtype_code(x), x.dim(), L.extents_arg(x.shape), _ptr(x), _ptr(out),
                                 _stream(x.device)), "afft_dft")
    return out


def dft(x):
    """DFT.hs:42-46: dftG (rootsOfUnity (shape v)) v."""
    return _run_dft(L.FORWARD, x)


def idft(x):
    """DFT.hs:51-60: dftG (inverseRootsOfUnity sh) v, divided by the innermost extent."""
    return _run_dft(L.INVERSE, x)


def dftG(roots, arr):
    x = _device_tensor(arr)
    r = _device_tensor(roots)
    if tuple(r.shape) != tuple(x.shape) or r.dtype != x.dtype:
        raise ValueError("the extent of the input and roots must match (DFT.hs:66)")
    out = torch.empty_like(x)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_dftG(_dtype_code(x), x.dim(), L.extents_arg(x.shape), _ptr(r), _ptr(x), _ptr(out),
                                  _stream(x.device)), "afft_dftG")
    return out


def dftGS(ix, roots, arr):
    x = _device_tensor(arr)
    r = _device_tensor(roots)
    if tuple(r.shape) != tuple(x.shape) or r.dtype != x.dtype:
        raise ValueError("the extent of the input and roots must match")
    out = torch.empty((), dtype=x.dtype, device=x.device)
    with torch.cuda.device(x.device):
        L.check(L.lib().afft_dftGS(_dtype_code(x), x.dim(), L.extents_arg(x.shape), L.extents_arg(ix), _ptr(r),
                                   _ptr(x), _ptr(out), _stream(x.device)), "afft_dftGS")
    return out


def _roots(shape, sign, dtype, device):
    n = int(shape[-1])
    i = np.arange(n, dtype=np.int64)
    # Roots.hs:26-51: k = 2 pi i / n; (cos k, -sin k) forward, (cos k, sin k) inverse
    k = 2 * math.pi * i / n if n else i.astype(np.float64)
    row = np.cos(k) + 1j * sign * np.sin(k)
    full = np.broadcast_to(row, tuple(int(s) for s in shape))
    return torch.from_numpy(np.ascontiguousarray(full)).to(dtype).to(device)


def rootsOfUnity(shape, dtype=torch.complex64, device="cuda"):
    return _roots(shape, -1, dtype, device)


def inverseRootsOfUnity(shape, dtype=torch.complex64, device="cuda"):
    return _roots(shape, +1, dtype, device)
