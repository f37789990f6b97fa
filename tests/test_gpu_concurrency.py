"""One shared plan executed at once from several host threads on several streams.

Every front-end call takes its plan from the shared cache, so two Accelerate streams calling
fft2D on one shape run the same plan concurrently -- the situation the reference leaves open
(LLVM/PTX/Plans.hs:58-63, "check that plans can be used concurrently"; the setStream/exec pair of
LLVM/PTX.hs:121-124).  The shapes here all have intermediate arrays: fft2D 4096^2 (rows -> work ->
two column passes), the 1 x 2^22 row (split six-step through the work array), 3 x 10007
(multi-kernel Bluestein through both auxiliary arrays) and the 10,000-transform ragged batch (long
classes through their workspaces, a side stream per execution).  Each thread transforms its own
seeded input several times on its own stream; its last output must match the FP64 oracle."""
import threading

import numpy as np
import pytest
import torch

import afft
import oracle as O
from afft import Mode
from conftest import tolerance
from test_gpu_ragged import check_rows, config_lengths

pytestmark = pytest.mark.gpu

THREADS = 4
REPEATS = 4


def run_threads(work):
    """Run work(i) in THREADS threads started together; re-raise the first failure."""
    errors = []
    start = threading.Barrier(THREADS)

    def body(i):
        try:
            start.wait()
            work(i)
        except Exception as e:  # reported below, with its thread
            errors.append((i, repr(e)))

    th = [threading.Thread(target=body, args=(i,)) for i in range(THREADS)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors


def same_plan_on_streams(gpu, call, inputs):
    """call(x) on each thread's input, REPEATS times on the thread's own stream; returns the last
    outputs, one per thread."""
    xs = [torch.from_numpy(x).to(gpu) for x in inputs]
    torch.cuda.synchronize()
    outs = [None] * THREADS

    def work(i):
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            for _ in range(REPEATS):
                outs[i] = call(xs[i])
        s.synchronize()

    run_threads(work)
    return [o.cpu().numpy() for o in outs]


@pytest.mark.parametrize("mode", [Mode.Forward, Mode.Inverse])
def test_fft2D_4096_same_plan_four_streams(gpu, mode):
    inputs = [O.normal(30 + i, (4096, 4096), np.complex64) for i in range(THREADS)]
    outs = same_plan_on_streams(gpu, lambda x: afft.fft2D(mode, x), inputs)
    om = O.Mode.Forward if mode is Mode.Forward else O.Mode.Inverse
    for i in range(THREADS):
        err = O.rel_l2(outs[i], O.fft2D(om, inputs[i]))
        assert err <= tolerance(False, 4096 * 4096), (i, err)


def test_row_2_22_same_plan_four_streams(gpu):
    n = 1 << 22
    inputs = [O.uniform(40 + i, (1, n), np.complex64) for i in range(THREADS)]
    outs = same_plan_on_streams(gpu, lambda x: afft.fft(Mode.Forward, x), inputs)
    for i in range(THREADS):
        err = O.rel_l2(outs[i], O.fft(O.Mode.Forward, inputs[i]))
        assert err <= tolerance(False, n), (i, err)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_bluestein_10007_same_plan_four_streams(gpu, dtype):
    inputs = [O.uniform(50 + i, (3, 10007), dtype) for i in range(THREADS)]
    outs = same_plan_on_streams(gpu, lambda x: afft.fft(Mode.Reverse, x), inputs)
    for i in range(THREADS):
        err = O.rel_l2(outs[i], O.fft(O.Mode.Reverse, inputs[i]))
        assert err <= tolerance(dtype == np.complex128, 10007), (i, err)


def test_explicit_plan_four_streams(gpu):
    # afft_exec on one explicit plan (a split 2^21 row, Complex Double) from four threads
    n = 1 << 21
    plan = afft.Plan((1, n), torch.complex128, "innermost")
    inputs = [O.uniform(60 + i, (1, n)) for i in range(THREADS)]
    outs = same_plan_on_streams(gpu, lambda x: plan.execute(x, -1, 1.0), inputs)
    for i in range(THREADS):
        err = O.rel_l2(outs[i], O.fft(O.Mode.Forward, inputs[i]))
        assert err <= tolerance(True, n), (i, err)
    plan.close()


def test_ragged_same_plan_four_streams(gpu):
    # the 10,000-transform configuration; threads 0-1 eager, threads 2-3 through captured graphs
    lengths = config_lengths()
    total = sum(lengths)
    plan = afft.RaggedPlan(lengths, torch.complex64)
    inputs = [O.uniform(70 + i, (total,), np.complex64) for i in range(THREADS)]
    xs = [torch.from_numpy(x).to(gpu) for x in inputs]
    outs = [torch.empty_like(x) for x in xs]
    torch.cuda.synchronize()

    def work(i):
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            for _ in range(REPEATS):
                plan.execute(Mode.Forward, xs[i], out=outs[i], graph=i >= 2)
        s.synchronize()

    run_threads(work)
    for i in range(THREADS):
        check_rows(outs[i], xs[i], lengths, Mode.Forward, torch.complex64)
    plan.close()
