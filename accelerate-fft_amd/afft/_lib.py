"""ctypes binding of libafft.so, the C ABI declared in include/afft.h.

The library is the only compute path: there is no CPU fallback.  If libafft.so has not been
built, or cannot be loaded, every call raises ``AfftError`` naming the missing file.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# AFFT_LIB names another build of the same library (the host-sanitised one, for instance)
LIB_PATH = os.environ.get("AFFT_LIB") or os.path.join(os.path.dirname(_HERE), "lib", "libafft.so")

SUCCESS = 0
C64 = 0
C128 = 1
FORWARD, REVERSE, INVERSE = 0, 1, 2
AXES_INNERMOST, AXES_ALL = 0, 1
REORDER_NONE, CENTRE, SHIFT, ISHIFT = 0, 1, 2, 3

_lib = None


class AfftError(RuntimeError):
    """A nonzero status from the C ABI (or the library missing)."""

    def __init__(self, status, what=""):
        self.status = status
        msg = strerror(status) if status >= 0 else "libafft.so not loadable"
        super().__init__(f"{what}: {msg} (status {status})" if what else f"{msg} (status {status})")


def _sig(f, res, *args):
    f.restype = res
    f.argtypes = list(args)


def lib():
    """Load libafft.so once and declare every entry point of include/afft.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise AfftError(-1, f"{LIB_PATH} is missing; build it with accelerate-fft_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    i, i64, vp, d = ctypes.c_int, ctypes.c_int64, ctypes.c_void_p, ctypes.c_double
    pi64 = ctypes.POINTER(ctypes.c_int64)
    pvp = ctypes.POINTER(ctypes.c_void_p)
    _sig(L.afft_strerror, ctypes.c_char_p, i)
    _sig(L.afft_version, i)
    _sig(L.afft_fft, i, i, i, i, pi64, vp, vp, vp)
    _sig(L.afft_fft1D, i, i, i, i64, vp, vp, vp)
    _sig(L.afft_fft2D, i, i, i, i64, i64, vp, vp, vp)
    _sig(L.afft_fft3D, i, i, i, i64, i64, i64, vp, vp, vp)
    _sig(L.afft_fft2D_many, i, i, i, i64, i64, i64, vp, vp, vp)
    _sig(L.afft_plan_create, i, pvp, i, i, pi64, i, i)
    _sig(L.afft_plan_cached, i, pvp, i, i, pi64, i, i)
    _sig(L.afft_exec, i, vp, i, vp, vp, d, vp)
    _sig(L.afft_exec_noalloc, i, vp, i, vp, vp, d, vp)
    _sig(L.afft_plan_destroy, i, vp)
    _sig(L.afft_cache_clear, i)
    _sig(L.afft_plan_info, i, vp, ctypes.POINTER(ctypes.c_int), pi64)
    _sig(L.afft_describe, i, i64, i, ctypes.POINTER(ctypes.c_int), i, ctypes.POINTER(ctypes.c_int),
         ctypes.POINTER(ctypes.c_int), pi64)
    _sig(L.afft_fft_vjp, i, i, i, i, i, pi64, vp, vp, vp)
    _sig(L.afft_ragged_plan_create, i, pvp, i, i64, pi64, i)
    _sig(L.afft_ragged_exec, i, vp, i, vp, vp, vp)
    _sig(L.afft_ragged_exec_graph, i, vp, i, vp, vp, vp)
    _sig(L.afft_ragged_plan_destroy, i, vp)
    _sig(L.afft_ragged_total, i, vp, pi64)
    _sig(L.afft_spectral_conv_fwd, i, i, i64, i64, i64, vp, vp, vp, vp)
    _sig(L.afft_spectral_conv_vjp, i, i, i64, i64, i64, vp, vp, vp, vp, vp, vp)
    _sig(L.afft_spectral_conv_fwd_save, i, i, i64, i64, i64, vp, vp, vp, vp, vp)
    _sig(L.afft_spectral_conv_vjp_saved, i, i, i64, i64, i64, vp, vp, vp, vp, vp, vp)
    _sig(L.afft_spectral_conv_vjp_residual, i, i, i64, i64, i64, vp, i, vp, vp, vp, ctypes.c_double, vp, vp, vp)
    _sig(L.afft_autotune, i, i, i, pi64, i, i)
    _sig(L.afft_ragged_autotune, i, i, i64, pi64, i)
    _sig(L.afft_tune_save, i, ctypes.c_char_p)
    _sig(L.afft_tune_load, i, ctypes.c_char_p)
    _sig(L.afft_reorder, i, i, i, i, pi64, vp, vp, vp)
    _sig(L.afft_fft_centred, i, i, i, i, pi64, i, i, vp, vp, vp)
    _sig(L.afft_dft, i, i, i, i, pi64, vp, vp, vp)
    _sig(L.afft_dftG, i, i, i, pi64, vp, vp, vp, vp)
    _sig(L.afft_dftGS, i, i, i, pi64, pi64, vp, vp, vp, vp)
    _lib = L
    return L


def strerror(status):
    try:
        return lib().afft_strerror(int(status)).decode()
    except AfftError:
        return "unknown status"


def check(status, what=""):
    if status != SUCCESS:
        raise AfftError(status, what)


def extents_arg(shape):
    return (ctypes.c_int64 * max(1, len(shape)))(*[int(s) for s in shape])
