"""Ragged batches (one persistent launch, afft_ragged_*) against the oracle, row by row."""
import math

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance

pytestmark = pytest.mark.gpu

OMODE = {Mode.Forward: O.Mode.Forward, Mode.Reverse: O.Mode.Reverse, Mode.Inverse: O.Mode.Inverse}
SIZE_SET = sorted({b << k for b in (1, 3, 5) for k in range(4, 14)} | {257, 1009, 4099})


def config_lengths(count=10000, seed=4):
    """log-uniform lengths in [16, 40960] snapped (in log distance) to the size set."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(math.log(16), math.log(40960), count)
    logs = np.log(np.array(SIZE_SET, dtype=np.float64))
    return [SIZE_SET[int(np.argmin(np.abs(logs - v)))] for v in u]


def check_rows(got, x, lengths, mode, dtype):
    got = got.cpu().numpy().astype(np.complex128)
    xr = x.cpu().numpy().astype(np.complex128)
    offs = np.concatenate([[0], np.cumsum(lengths)])
    by = {}
    for i, n in enumerate(lengths):
        by.setdefault(n, []).append(i)
    worst = 0.0
    for n, idx in by.items():
        if n == 0:
            continue
        xs = np.stack([xr[offs[i]:offs[i] + n] for i in idx])
        gs = np.stack([got[offs[i]:offs[i] + n] for i in idx])
        want = O.fft(OMODE[mode], xs)
        for r in range(len(idx)):
            err = O.rel_l2(gs[r], want[r])
            worst = max(worst, err / tolerance(dtype == torch.complex128, n))
            assert err <= tolerance(dtype == torch.complex128, n), (n, mode, err)
    return worst


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128])
def test_ragged_mixed_lengths(gpu, dtype):
    rng = np.random.default_rng(1)
    extra = [1, 2, 3, 7, 15, 16, 100, 243, 1000, 4096, 5000, 10007, 24576, 40960]
    lengths = list(rng.choice(SIZE_SET, 300)) + extra
    lengths = [int(n) for n in rng.permutation(lengths)]
    total = sum(lengths)
    x = torch.from_numpy(O.uniform(3, (total,))).to(dtype).to(gpu)
    plan = afft.RaggedPlan(lengths, dtype)
    for mode in Mode:
        y = plan.execute(mode, x)
        check_rows(y, x, lengths, mode, dtype)
    plan.close()


def test_ragged_in_place_and_empty_rows(gpu):
    lengths = [64, 0, 4099, 0, 3, 20480, 1]
    total = sum(lengths)
    x = torch.from_numpy(O.uniform(5, (total,))).to(torch.complex64).to(gpu)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    y = x.clone()
    out = plan.execute(Mode.Forward, y, out=y)
    assert out.data_ptr() == y.data_ptr()
    check_rows(y, x, lengths, Mode.Forward, torch.complex64)
    # every transform reads and writes only its own rows: in place through the graph too
    z = x.clone()
    for _ in range(2):
        z.copy_(x)
        plan.execute(Mode.Inverse, z, out=z, graph=True)
        torch.cuda.synchronize()
        check_rows(z, x, lengths, Mode.Inverse, torch.complex64)
    plan.close()


def test_ragged_graph_replays(gpu):
    lengths = config_lengths(800, seed=9)
    total = sum(lengths)
    x = torch.from_numpy(O.uniform(6, (total,))).to(torch.complex64).to(gpu)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    eager = plan.execute(Mode.Inverse, x)
    out = torch.empty_like(x)
    for rep in range(4):
        out.zero_()
        plan.execute(Mode.Inverse, x, out=out, graph=True)
        torch.cuda.synchronize()
        # every replay, with eager work between replays, matches the eager launch
        assert torch.allclose(out, eager, rtol=0, atol=0), rep
        _ = plan.execute(Mode.Forward, x)  # eager work between replays
    check_rows(out, x, lengths, Mode.Inverse, torch.complex64)
    plan.close()


def test_ragged_full_configuration(gpu):
    # 10,000 Complex Float transforms, lengths log-uniform over the size set (seed 4)
    lengths = config_lengths()
    total = sum(lengths)
    assert 4e7 < total < 7e7
    x = torch.from_numpy(O.uniform(4, (total,))).to(torch.complex64).to(gpu)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    y = plan.execute(Mode.Forward, x)
    check_rows(y, x, lengths, Mode.Forward, torch.complex64)
    z = plan.execute(Mode.Inverse, y)
    assert O.rel_l2(z.cpu().numpy(), x.cpu().numpy()) < 2e-6
    plan.close()


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Inverse])
def test_ragged_compiled_and_table_layouts(gpu, monkeypatch, mode):
    # every length of the persistent kernel's compiled list (Complex Float), the chirp-z classes
    # whose padded lengths are in it (7 -> 16, 257 -> 640, 1009 -> 2048), partial items (counts
    # that do not fill a workgroup) and lengths outside the list (15, 100, 243): the compiled
    # layouts and the pass-table layouts (AFFT_RAGGED_NO_CT=1) each against the oracle
    rng = np.random.default_rng(12)
    short = [n for n in SIZE_SET if n <= 4096]
    lengths = [n for n in short + [7, 15, 100, 243] for _ in range(int(rng.integers(1, 40)))]
    lengths = [int(n) for n in rng.permutation(lengths)]
    total = sum(lengths)
    x = torch.from_numpy(O.uniform(13, (total,))).to(torch.complex64).to(gpu)
    outs = []
    for no_ct in ("0", "1"):
        monkeypatch.setenv("AFFT_RAGGED_NO_CT", no_ct)
        plan = afft.RaggedPlan(lengths, torch.complex64)
        y = plan.execute(mode, x)
        check_rows(y, x, lengths, mode, torch.complex64)
        outs.append(y)
        plan.close()
    assert O.rel_l2(outs[0].cpu().numpy(), outs[1].cpu().numpy()) <= tolerance(False, 4096)


def test_ragged_items_beyond_the_grid(gpu):
    # more work items than the persistent grid holds (up to four rounds of resident workgroups,
    # 3072 on 256 CUs): 3500 transforms of 4096 points (one item each) and a mixture of shorter
    # classes, so workgroups walk several items of different classes in turn
    rng = np.random.default_rng(21)
    short = [n for n in config_lengths(600, seed=21) if n <= 4096]
    lengths = [4096] * 3500 + [2048] * 300 + short
    lengths = [int(n) for n in rng.permutation(lengths)]
    total = sum(lengths)
    x = torch.from_numpy(O.uniform(22, (total,))).to(torch.complex64).to(gpu)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    for mode in (Mode.Forward, Mode.Inverse):
        y = plan.execute(mode, x)
        check_rows(y, x, lengths, mode, torch.complex64)
    plan.close()
