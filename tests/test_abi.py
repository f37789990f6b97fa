"""The C ABI without a GPU: the library loads, exports every symbol include/afft.h declares, and
its host-only planning and argument checks behave (no compute call is made here)."""
import ctypes
import os
import re

import pytest

import afft
from afft import _lib as L

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "afft.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(afft_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = L.lib()
    names = declared_symbols()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name


def test_strerror_and_version():
    lib = L.lib()
    for code in range(0, 10):
        assert lib.afft_strerror(code)
    assert lib.afft_strerror(0) == b"success"
    assert lib.afft_strerror(12345) == b"unknown status"
    assert lib.afft_version() >= 10000


def is_5smooth(n):
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_describe_layouts(dtype):
    import torch
    dt = torch.complex64 if dtype == "c64" else torch.complex128
    limit = 16384 if dtype == "c64" else 8192
    for n in list(range(1, 1025)) + [2048, 3072, 4096, 5120, 6144, 8192, 10240, 12288, 16384]:
        if n > limit and is_5smooth(n):
            continue
        radices, E, blue = afft.describe(n, dt)
        M = blue if blue else n
        prod = 1
        for r in radices:
            assert r in (2, 3, 4, 5, 8, 16)
            assert E % r == 0, (n, E, r)
            prod *= r
        assert prod == M, (n, radices)
        assert M % E == 0
        if blue:
            assert not is_5smooth(n)
            assert blue >= 2 * n - 1 and is_5smooth(blue)
        else:
            assert is_5smooth(n)


def test_describe_known_layouts():
    import torch
    assert afft.describe(4096, torch.complex64) == ([16, 16, 16], 16, 0)
    assert afft.describe(1024, torch.complex128)[0] == [16, 16, 4]
    assert afft.describe(1, torch.complex64) == ([], 1, 0)
    # primes of the ragged-batch set go to Bluestein
    assert afft.describe(257, torch.complex64)[2] == 640
    assert afft.describe(1009, torch.complex64)[2] == 2048
    assert afft.describe(4099, torch.complex64)[2] == 10240


def test_describe_rejects():
    lib = L.lib()
    r = (ctypes.c_int * 4)()
    a, b = ctypes.c_int(), ctypes.c_int()
    c = ctypes.c_int64()
    assert lib.afft_describe(0, 0, r, 4, a, b, c) == 1          # invalid length
    assert lib.afft_describe(16, 7, r, 4, a, b, c) == 3         # bad dtype
    assert lib.afft_describe(1 << 15, 0, r, 4, a, b, c) == 4    # beyond one workgroup (four-step)
    assert lib.afft_describe(1 << 14, 1, r, 4, a, b, c) == 4


def test_argument_checks_before_any_device_work():
    lib = L.lib()
    h = ctypes.c_void_p()
    ext = L.extents_arg([16])
    assert lib.afft_plan_create(None, 0, 1, ext, 0, 0) == 1
    assert lib.afft_plan_create(ctypes.byref(h), 5, 1, ext, 0, 0) == 3
    assert lib.afft_plan_create(ctypes.byref(h), 0, 0, ext, 0, 0) == 2
    assert lib.afft_plan_create(ctypes.byref(h), 0, 9, L.extents_arg([2] * 9), 0, 0) == 2
    assert lib.afft_plan_create(ctypes.byref(h), 0, 1, L.extents_arg([-4]), 0, 0) == 1
    assert lib.afft_plan_create(ctypes.byref(h), 0, 1, ext, 7, 0) == 1
    assert lib.afft_exec(None, -1, None, None, 1.0, None) == 1
    assert lib.afft_fft(9, 0, 1, ext, None, None, None) == 1          # bad mode
    assert lib.afft_fft(0, 0, 1, ext, None, None, None) in (1, 7)      # null pointers / no device
    # empty arrays are a no-op that touches nothing
    assert lib.afft_fft(0, 0, 2, L.extents_arg([0, 8]), None, None, None) == 0
    assert lib.afft_fft2D(0, 1, 0, 5, None, None, None) == 0
    assert lib.afft_fft_vjp(2, 0, 0, 1, ext, None, None, None) == 2   # fft2D needs rank 2
    assert lib.afft_fft_vjp(9, 0, 0, 1, ext, None, None, None) == 1


def test_tune_table_file_handling(tmp_path):
    lib = L.lib()
    # save writes the header even for an empty table
    p = tmp_path / "t.txt"
    assert lib.afft_tune_save(str(p).encode()) == 0
    assert p.read_text().startswith("afft-tune 1")
    # a corrupt file is refused with an I/O status, a missing one too
    bad = tmp_path / "bad.txt"
    bad.write_text("not a tune file\n")
    assert lib.afft_tune_load(str(bad).encode()) == 8
    assert lib.afft_tune_load(str(tmp_path / "missing.txt").encode()) == 8
    # malformed and stale lines are skipped without error
    ok = tmp_path / "ok.txt"
    ok.write_text("afft-tune 1\n"
                  "gfx950 70226015 rows 0 4096 1 190.0\n"
                  "gfx950 70226015 rows 0 4096 999 1.0\n"
                  "garbage line\n"
                  "gfx950 70226015 rows 7 4096 0 1.0\n"
                  "gfx950 70226015 strided 0 512 16016 44.0\n"
                  "gfx950 70226015 strided 0 512 16007 44.0\n"
                  "gfx950 70226015 regcol 0 64 4 40.0\n"
                  "gfx950 70226015 regcol 0 64 3 40.0\n"
                  "gfx950 70226015 ragged 0 1000 8 5.0\n"
                  "gfx950 70226015 ragged 0 1000 999 5.0\n"
                  "gfx950 70226015 nonsense 0 1000 8 5.0\n")
    assert lib.afft_tune_load(str(ok).encode()) == 0
    out = tmp_path / "out.txt"
    assert lib.afft_tune_save(str(out).encode()) == 0
    lines = out.read_text().splitlines()
    assert lines[0] == "afft-tune 1"
    assert any(l.split()[:6] == ["gfx950", "70226015", "rows", "0", "4096", "1"] for l in lines[1:])
    assert any(l.split()[2:6] == ["strided", "0", "512", "16016"] for l in lines[1:])
    assert any(l.split()[2:6] == ["regcol", "0", "64", "4"] for l in lines[1:])
    assert any(l.split()[2:6] == ["ragged", "0", "1000", "8"] for l in lines[1:])
    assert not any(" 999 " in l or " 16007 " in l or "regcol 0 64 3 " in l or "nonsense" in l for l in lines)
