"""CPU oracle for the Data.Array.Accelerate.Math.FFT hot path (test infrastructure only).

Only the test-suite, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, and only as the checker (or as the CPU path timed beside the GPU).  The shipped library
never calls it.

It wraps ``libafft_oracle.so`` (built from ``afft_oracle.c`` by ``oracle/Makefile``) and restates
the reference's front-end semantics on numpy arrays:

* ``Mode`` / ``sign_of_mode``  -- src/Data/Array/Accelerate/Math/FFT/Mode.hs:15-26
* ``fft``   -- innermost axis, outer axes are a batch, Inverse divides by the innermost extent
               (src/Data/Array/Accelerate/Math/FFT.hs:63-84; LLVM/Native.hs:49-72 axis [rank-1])
* ``fft1D`` -- src/Data/Array/Accelerate/Math/FFT.hs:92-111 (Inverse divides by the length)
* ``fft2D`` / ``fft3D`` -- all axes, Inverse divides by the total size
               (FFT.hs:119-173; LLVM/Native.hs:84-102 axes [0,1] and [0,1,2])
* ``dft_brute`` -- the O(N^2) definition of src/Data/Array/Accelerate/Math/DFT.hs:68-111 with
               the roots of DFT/Roots.hs:26-51, in long double.
* ``centre`` / ``shift`` / ``ishift`` -- src/Data/Array/Accelerate/Math/DFT/Centre.hs:36-164, as
               an index-by-index restatement of the generate / backpermute definitions.
* ``dftG`` -- src/Data/Array/Accelerate/Math/DFT.hs:63-92, the extended-array sum over every
               element, in FP64 with the index product reduced in integers.

Arrays are numpy ``complex128`` (FP64 throughout) in Accelerate's row-major order, innermost
index fastest; ``complex64`` inputs are promoted, so the oracle consumes the rounded values.
"""
import ctypes
import enum
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libafft_oracle.so")
_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        i64, dp, fp = ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p
        L.orc_gen_uniform.argtypes = [ctypes.c_uint64, i64, dp]
        L.orc_gen_normal.argtypes = [ctypes.c_uint64, i64, dp]
        L.orc_dft_brute.argtypes = [i64, ctypes.c_int, dp, dp]
        L.orc_fft_rows_f64.argtypes = [i64, ctypes.c_int, i64, ctypes.c_double, dp, dp, ctypes.c_int]
        L.orc_fft_rows_f32.argtypes = [i64, ctypes.c_int, i64, ctypes.c_double, fp, fp, ctypes.c_int]
        L.orc_fft_nd_f64.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, dp, dp, ctypes.c_int]
        for f in (L.orc_dft_brute, L.orc_fft_rows_f64, L.orc_fft_rows_f32, L.orc_fft_nd_f64):
            f.restype = ctypes.c_int
        _lib = L
    return _lib


class Mode(enum.Enum):
    """Transform direction, as Mode.hs:15-19."""
    Forward = 0
    Reverse = 1
    Inverse = 2


def sign_of_mode(mode):
    """Mode.hs:21-26: Forward -1, Reverse +1, Inverse +1."""
    return -1 if Mode(mode) is Mode.Forward else 1


def _nthreads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _check(status, what):
    if status != 0:
        raise RuntimeError(f"oracle {what} failed with status {status}")


# ---------------------------------------------------------------------------------------------
# seeded inputs

def uniform(seed, shape, dtype=np.complex128):
    """re/im i.i.d. uniform[-1,1) from the counter-based SplitMix64 stream `seed`
    (re of element i from counter 2i, im from 2i+1), rounded to `dtype`."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape, dtype=np.int64)) if shape else 1
    buf = np.empty(2 * n, dtype=np.float64)
    lib().orc_gen_uniform(seed, 2 * n, buf.ctypes.data)
    return buf.view(np.complex128).reshape(shape).astype(dtype)


def normal(seed, shape, dtype=np.complex128):
    """re/im i.i.d. standard normal (Box-Muller over counter pairs), rounded to `dtype`."""
    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape, dtype=np.int64)) if shape else 1
    buf = np.empty(2 * n, dtype=np.float64)
    lib().orc_gen_normal(seed, 2 * n, buf.ctypes.data)
    return buf.view(np.complex128).reshape(shape).astype(dtype)


# ---------------------------------------------------------------------------------------------
# transforms

def dft_brute(x, sign):
    """O(N^2) DFT of a 1D vector in long double (DFT.hs:68-92), returned as complex128."""
    x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    out = np.empty_like(x)
    _check(lib().orc_dft_brute(x.size, sign, x.ctypes.data, out.ctypes.data), "dft_brute")
    return out


def fft_rows(x, sign, scale=1.0, nthreads=None):
    """Unnormalised DFT of every contiguous innermost row, times `scale`, in FP64."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    if x.size == 0:
        return out
    n = x.shape[-1]
    _check(lib().orc_fft_rows_f64(n, sign, x.size // n, float(scale), x.ctypes.data,
                                  out.ctypes.data, nthreads or _nthreads()), "fft_rows")
    return out


def fft_rows_f32(x, sign, scale=1.0, nthreads=None):
    """The same recursion in float arithmetic (twiddles from long double); used only to time
    a single-precision CPU path beside the GPU, never as the accuracy oracle."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    out = np.empty_like(x)
    if x.size == 0:
        return out
    n = x.shape[-1]
    _check(lib().orc_fft_rows_f32(n, sign, x.size // n, float(scale), x.ctypes.data,
                                  out.ctypes.data, nthreads or _nthreads()), "fft_rows_f32")
    return out


def fft_axes(x, axes, sign, scale=1.0, nthreads=None):
    """Unnormalised DFT along `axes` (0 = outermost), times `scale`, in FP64."""
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.empty_like(x)
    if x.size == 0:
        return out
    rank = x.ndim
    ext = (ctypes.c_int64 * rank)(*x.shape)
    mask = 0
    for a in axes:
        mask |= 1 << (a % rank)
    _check(lib().orc_fft_nd_f64(rank, ext, mask, sign, float(scale), x.ctypes.data,
                                out.ctypes.data, nthreads or _nthreads()), "fft_nd")
    return out


def _scale(mode, divisor):
    return 1.0 / divisor if Mode(mode) is Mode.Inverse and divisor > 0 else 1.0


def fft(mode, arr):
    """FFT.hs:63-84: innermost axis; Inverse divides by the innermost extent (FFT.hs:69)."""
    arr = np.asarray(arr)
    return fft_rows(arr, sign_of_mode(mode), _scale(mode, arr.shape[-1]))


def fft1D(mode, arr):
    """FFT.hs:92-111: a vector; Inverse divides by its length (FFT.hs:98)."""
    arr = np.asarray(arr)
    if arr.ndim != 1:
        raise ValueError("fft1D takes a rank-1 array")
    return fft_rows(arr, sign_of_mode(mode), _scale(mode, arr.shape[0]))


def fft2D(mode, arr):
    """FFT.hs:119-142: both axes; Inverse divides by the total size (FFT.hs:125)."""
    arr = np.asarray(arr)
    if arr.ndim != 2:
        raise ValueError("fft2D takes a rank-2 array")
    return fft_axes(arr, (0, 1), sign_of_mode(mode), _scale(mode, arr.size))


def fft3D(mode, arr):
    """FFT.hs:150-173: all three axes; Inverse divides by the total size (FFT.hs:155)."""
    arr = np.asarray(arr)
    if arr.ndim != 3:
        raise ValueError("fft3D takes a rank-3 array")
    return fft_axes(arr, (0, 1, 2), sign_of_mode(mode), _scale(mode, arr.size))


def rel_l2(got, want):
    """Relative L2 error ||got - want|| / ||want|| in FP64 (0 when both are zero)."""
    got = np.asarray(got, dtype=np.complex128)
    want = np.asarray(want, dtype=np.complex128)
    den = np.linalg.norm(want.ravel())
    num = np.linalg.norm((got - want).ravel())
    if den == 0.0:
        return float(num)
    return float(num / den)


# ----------------------------------------------------------------------------------------------
# DFT/Centre.hs, restated element by element from the index functions (not with numpy's
# fftshift, which the tests use as an independent check).

def centre(arr):
    """centre1D/2D/3D (Centre.hs:36-65): x[ix] * ((-1) ** sum(ix) :+ 0)."""
    x = np.asarray(arr, dtype=np.complex128)
    out = np.empty_like(x)
    for ix in np.ndindex(*x.shape):
        out[ix] = x[ix] * (-1.0) ** sum(ix)
    return out


def _roll_back(arr, shift_of):
    """backpermute sh p arr with p(ix)_a = (ix_a + shift_of(n_a)) `rem` n_a."""
    x = np.asarray(arr, dtype=np.complex128)
    out = np.empty_like(x)
    for ix in np.ndindex(*x.shape):
        src = tuple((i + shift_of(n)) % n for i, n in zip(ix, x.shape))
        out[ix] = x[src]
    return out


def shift(arr):
    """shift1D/2D/3D (Centre.hs:70-82, 104-123): shift_a = n_a `quot` 2 + boolToInt (odd n_a)."""
    return _roll_back(arr, lambda n: n // 2 + (n % 2))


def ishift(arr):
    """ishift1D/2D/3D (Centre.hs:84-101, 125-164): shift_a = n_a `quot` 2."""
    return _roll_back(arr, lambda n: n // 2)


# ----------------------------------------------------------------------------------------------
# Math/DFT.hs

def dftG(roots, arr):
    """DFT.hs:63-92: out[ix] = sum over every linear index i of arr[i] * roots[sh_i :. (k*n_i) mod l],
    (sh_i :. n_i) = fromIndex (shape arr) i, k = indexHead ix."""
    x = np.asarray(arr, dtype=np.complex128)
    r = np.asarray(roots, dtype=np.complex128)
    l = x.shape[-1]
    xf = x.reshape(-1, l)
    rf = r.reshape(-1, l)
    n = np.arange(l, dtype=np.int64)
    vals = np.empty(l, dtype=np.complex128)
    for k in range(l):
        e = (k * n) % l
        vals[k] = np.sum(xf * rf[:, e])
    return np.broadcast_to(vals, x.shape).copy()

