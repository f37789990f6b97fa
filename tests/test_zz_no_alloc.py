"""afft_exec_noalloc: afft_exec with its intermediate arrays in the library's fixed per-GPU scratch
instead of a leased pool set, so the call allocates nothing.  Checked through the C ABI against
afft_exec on the same plan and input (identical kernels: identical bits) and against the FP64
oracle, on two calls in a row, the second with a smaller input."""
import ctypes

import numpy as np
import pytest
import torch

import afft
import oracle as O
from conftest import tolerance

pytestmark = pytest.mark.gpu

SCRATCH_TOO_SMALL = 10


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _noalloc(plan, direction, x, out, scale):
    return afft._lib.lib().afft_exec_noalloc(plan._h, direction, _ptr(x), _ptr(out), float(scale), _stream())


def test_noalloc_matches_exec_on_two_calls(gpu):
    # 1: a 2^22-point Complex Float row, Forward (split six-step: 32 MiB work array)
    # 2: 4 rows of 10007 Complex Double, Inverse (multi-kernel Bluestein: the auxiliary arrays)
    cases = [((1, 1 << 22), torch.complex64, -1, 1.0, O.Mode.Forward),
             ((4, 10007), torch.complex128, +1, 1.0 / 10007, O.Mode.Inverse)]
    for i, (shape, dtype, direction, scale, omode) in enumerate(cases):
        print(f"afft_exec_noalloc case {i}: shape {shape} {dtype}", flush=True)
        plan = afft.Plan(shape, dtype, "innermost")
        assert plan.info()["launches"] > 1
        xh = O.uniform(80 + i, shape)
        x = torch.from_numpy(xh).to(dtype).to(gpu)
        ref = plan.execute(x, direction, scale)  # afft_exec: leases from the plan's pool
        out = torch.empty_like(x)
        torch.cuda.synchronize()
        free_before = torch.cuda.mem_get_info()[0]
        assert _noalloc(plan, direction, x, out, scale) == 0
        torch.cuda.synchronize()
        if i > 0:  # the first call may load the library's scratch; later calls allocate nothing
            assert torch.cuda.mem_get_info()[0] == free_before
        assert torch.equal(out, ref), f"case {i}: afft_exec_noalloc differs from afft_exec"
        err = O.rel_l2(out.cpu().numpy(), O.fft(omode, x.cpu().numpy()))
        assert err <= tolerance(dtype == torch.complex128, shape[-1]), (i, err)
        plan.close()


def test_noalloc_without_intermediates_and_too_large(gpu):
    # a single-pass plan needs no scratch at all
    plan = afft.Plan((64, 4096), torch.complex64, "innermost")
    x = torch.from_numpy(O.uniform(90, (64, 4096))).to(torch.complex64).to(gpu)
    out = torch.empty_like(x)
    assert _noalloc(plan, -1, x, out, 1.0) == 0
    torch.cuda.synchronize()
    assert torch.equal(out, plan.execute(x, -1, 1.0))
    plan.close()
    # fft2D 4096^2 Complex Float needs a 128 MiB work array, beyond the 64 MiB scratch: refused
    # before any launch
    big = afft.Plan((4096, 4096), torch.complex64, "all")
    y = torch.zeros((4096, 4096), dtype=torch.complex64, device=gpu)
    z = torch.empty_like(y)
    assert _noalloc(big, -1, y, z, 1.0) == SCRATCH_TOO_SMALL
    big.close()
