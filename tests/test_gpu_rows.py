"""The persistent row kernels: every compiled variant of every length, with batches that run the
persistent loop several times and end in a partial tile, and the autotuner round trip."""
import ctypes

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu

LENGTHS = [(torch.complex64, n) for n in (256, 512, 1024, 2048, 4096, 8192, 5120, 6144, 10240, 12288,
                                          160, 320, 384, 640, 768, 1536)] + \
          [(torch.complex128, n) for n in (256, 512, 1024, 2048, 4096)]


@pytest.mark.parametrize("dtype,n", LENGTHS)
def test_every_row_variant(gpu, dtype, n, monkeypatch):
    rows = max(3000, (1 << 23) // n) + 5  # several persistent iterations and a partial tile
    x = torch.from_numpy(O.uniform(n, (rows, n))).to(dtype).to(gpu)
    xr = x.cpu().numpy().astype(np.complex128)
    pick = np.r_[0:40, rows - 40:rows]
    want = {m: O.fft(m, xr[pick]) for m in (O.Mode.Forward, O.Mode.Reverse)}
    tol = tolerance(dtype == torch.complex128, n)
    for v in range(4):
        monkeypatch.setenv("AFFT_ROW_VARIANT", str(v))
        p = afft.Plan((rows, n), dtype, "innermost")
        assert p.info()["launches"] == 1
        y = p.execute(x, -1, 1.0)
        assert O.rel_l2(y.cpu().numpy()[pick], want[O.Mode.Forward]) <= tol, v
        z = p.execute(x, +1, 0.5)
        assert O.rel_l2(z.cpu().numpy()[pick], 0.5 * want[O.Mode.Reverse]) <= tol, v
        p.close()


def test_autotune_round_trip(gpu, tmp_path):
    lib = afft._lib.lib()
    ext = afft._lib.extents_arg([64, 4096])
    assert lib.afft_autotune(0, 2, ext, 0, 0) == 0
    path = tmp_path / "tune.txt"
    assert lib.afft_tune_save(str(path).encode()) == 0
    lines = path.read_text().splitlines()
    assert lines[0] == "afft-tune 1"
    entry = [l.split() for l in lines[1:] if l.split()[2:5] == ["rows", "0", "4096"]]
    assert entry and entry[0][0].startswith("gfx")
    assert lib.afft_tune_load(str(path).encode()) == 0
    x = torch.from_numpy(O.uniform(3, (64, 4096))).to(torch.complex64).to(gpu)
    y = afft.fft(Mode.Forward, x)
    assert O.rel_l2(y.cpu().numpy(), O.fft(O.Mode.Forward, x.cpu().numpy())) <= tolerance(False, 4096)
