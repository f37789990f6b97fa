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
                                          160, 320, 384, 640, 768, 1536,
                                          16384, 20480)] + \
          [(torch.complex128, n) for n in (256, 512, 1024, 2048, 4096)]


@pytest.mark.parametrize("dtype,n", LENGTHS)
def test_every_row_variant(gpu, dtype, n, monkeypatch):
    # several persistent iterations and a partial tile (rows beyond 16384 points take the row
    # kernel for up to 4 rows per CU, a six-step beyond)
    rows = (1000 if n > 16384 else max(3000, (1 << 23) // n)) + 5
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
    # a row length, a strided tile (the 512-point columns of a batched 2D) and a register column
    # pass (the split axis of fft2D 4096^2): every decision point of each plan is timed
    assert lib.afft_autotune(0, 2, afft._lib.extents_arg([64, 4096]), 0, 0) == 0
    assert lib.afft_autotune(0, 3, afft._lib.extents_arg([16, 512, 512]), 2, 0) == 0
    assert lib.afft_autotune(0, 2, afft._lib.extents_arg([4096, 4096]), 1, 0) == 0
    # a ragged batch: every class timed alone (persistent-kernel classes and a long class)
    lengths = [64] * 300 + [1000] * 60 + [4096] * 30 + [257] * 40 + [20480] * 8
    arr = (ctypes.c_int64 * len(lengths))(*lengths)
    assert lib.afft_ragged_autotune(0, len(lengths), arr, 0) == 0
    path = tmp_path / "tune.txt"
    assert lib.afft_tune_save(str(path).encode()) == 0
    lines = path.read_text().splitlines()
    assert lines[0] == "afft-tune 1"
    rows = [l.split() for l in lines[1:]]
    kinds = {r[2] for r in rows}
    assert {"rows", "strided", "regcol", "ragged", "ragged_split"} <= kinds, kinds
    assert any(r[2:5] == ["rows", "0", "4096"] for r in rows)
    assert any(r[2:5] == ["strided", "0", "512"] for r in rows)
    assert any(r[2:5] == ["regcol", "0", "64"] for r in rows)
    assert any(r[2] == "ragged" and r[4] in ("64", "1000", "4096", "257") for r in rows)
    # the whole batch's side-stream split: the default (its one long class, 20480 x 8, aside) or
    # none aside
    split = [r for r in rows if r[2] == "ragged_split"]
    assert len(split) == 1 and int(split[0][5]) in (0, 1 << 22), split
    assert all(r[0].startswith("gfx") for r in rows)
    # a stale and a corrupt entry next to the good ones are ignored on load
    arch, hv = rows[0][0], rows[0][1]
    text = path.read_text() + f"{arch} {hv} strided 0 512 16007 1.0\n{arch} {hv} ragged 0 64 999 1.0\n" \
                              f"{arch} {hv} regcol 0 64 3 1.0\nnot an entry\n"
    path.write_text(text)
    assert lib.afft_tune_load(str(path).encode()) == 0
    again = tmp_path / "again.txt"
    assert lib.afft_tune_save(str(again).encode()) == 0
    kept = again.read_text()
    assert " 16007 " not in kept and " 999 " not in kept and "regcol 0 64 3 " not in kept
    # plans built with the tuned choices are still right
    x = torch.from_numpy(O.uniform(3, (16, 512, 512))).to(torch.complex64).to(gpu)
    y = afft.fft2D_many(Mode.Forward, x)
    assert O.rel_l2(y.cpu().numpy(), O.fft_axes(x.cpu().numpy(), (1, 2), -1)) <= tolerance(False, 512 * 512)
    z = torch.from_numpy(O.uniform(4, (1024, 4096))).to(torch.complex64).to(gpu)
    w = afft.fft2D(Mode.Inverse, z)
    assert O.rel_l2(w.cpu().numpy(), O.fft2D(O.Mode.Inverse, z.cpu().numpy())) <= tolerance(False, 1024 * 4096)
    total = sum(lengths)
    r = torch.from_numpy(O.uniform(5, (total,))).to(torch.complex64).to(gpu)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    out = plan.execute(Mode.Forward, r).cpu().numpy()
    xr = r.cpu().numpy()
    off = 0
    for n in (64, 1000, 4096, 257, 20480):
        cnt = lengths.count(n)
        got = out[off:off + cnt * n].reshape(cnt, n)
        assert O.rel_l2(got, O.fft(O.Mode.Forward, xr[off:off + cnt * n].reshape(cnt, n))) <= tolerance(False, n), n
        off += cnt * n
    plan.close()


@pytest.mark.parametrize("n", [20480, 24576, 40960])
def test_transposed_rows_at_every_output_alignment(gpu, n):
    # a six-step plan 64 x (n / 64) whose last pass stores transposed, in tiles of rows rotated by
    # the output's alignment.  The rotation only moves rows between workgroups, so an output at
    # any element offset holds the aligned result bit for bit, and nothing around it is written.
    shape = (3, n)
    plan = afft.Plan(shape, torch.complex64, "innermost")
    xh = O.uniform(31, shape)
    x = torch.from_numpy(xh).to(torch.complex64).to(gpu)
    ref = plan.execute(x, -1, 1.0)
    assert O.rel_l2(ref.cpu().numpy(), O.fft(O.Mode.Forward, x.cpu().numpy())) <= tolerance(False, n)
    total = x.numel()
    buf = torch.empty(total + 40, dtype=torch.complex64, device=gpu)
    for off in (1, 3, 8, 15, 16, 17, 29):
        buf.fill_(complex(float("nan"), float("nan")))
        out = buf[off:off + total].view(shape)
        plan.execute(x, -1, 1.0, out=out)
        assert torch.equal(out, ref), off
        assert torch.isnan(buf[:off]).all() and torch.isnan(buf[off + total:]).all(), off
    plan.close()


@pytest.mark.parametrize("dtype,n,rows", [
    (torch.complex64, 256, 262144 + 7),    # 512 MiB: one tile per workgroup
    (torch.complex64, 1024, 40000 + 3),    # 312 MiB: one tile per workgroup, the default variant 1
    (torch.complex64, 2048, 9000 + 1),     # 140 MiB: four resident rounds
    (torch.complex128, 512, 40000 + 5),    # 312 MiB
    (torch.complex128, 2048, 5000 + 3),    # 156 MiB: four resident rounds
])
def test_rows_grid_beyond_resident(gpu, dtype, n, rows):
    # large batches launch more row workgroups than are resident (four rounds beyond 128 MiB, one
    # tile per workgroup beyond 256 MiB), ending in a partial tile; sampled rows against the oracle
    x = torch.from_numpy(O.uniform(41, (rows, n))).to(dtype).to(gpu)
    pick = np.r_[0:24, rows // 2 - 12:rows // 2 + 12, rows - 24:rows]
    xr = x[pick].cpu().numpy().astype(np.complex128)
    tol = tolerance(dtype == torch.complex128, n)
    p = afft.Plan((rows, n), dtype, "innermost")
    assert p.info()["launches"] == 1
    y = p.execute(x, -1, 1.0)
    assert O.rel_l2(y[pick].cpu().numpy(), O.fft(O.Mode.Forward, xr)) <= tol
    z = p.execute(x, +1, 1.0 / n)
    assert O.rel_l2(z[pick].cpu().numpy(), O.fft(O.Mode.Inverse, xr)) <= tol
    p.close()
