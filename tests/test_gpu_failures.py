"""Failure handling on the GPU, forced from the host: every device allocation a plan creation makes
is failed in turn (AFFT_TEST_FAIL_ALLOC=k fails the k-th), and each time the call must return
AFFT_ERROR_ALLOC, hand back no plan and leave the device's free memory as it found it."""
import ctypes

import pytest
import torch

import afft
from afft import _lib as L

pytestmark = pytest.mark.gpu

ALLOC = 5


def _create(shape, dtype, axes):
    h = ctypes.c_void_p()
    st = L.lib().afft_plan_create(ctypes.byref(h), dtype, len(shape), L.extents_arg(shape), axes, 0)
    return st, h


def _create_ragged(lengths, dtype):
    h = ctypes.c_void_p()
    arr = (ctypes.c_int64 * len(lengths))(*lengths)
    st = L.lib().afft_ragged_plan_create(ctypes.byref(h), dtype, len(lengths), arr, 0)
    return st, h


def _fail_each_allocation(monkeypatch, create, destroy):
    # one full creation first, so that the runtime's own pools are warm before free memory is read
    st, h = create()
    assert st == 0
    destroy(h)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    k = 1
    while True:
        monkeypatch.setenv("AFFT_TEST_FAIL_ALLOC", str(k))
        st, h = create()
        if st == 0:
            break
        assert st == ALLOC, (k, st)
        assert not h.value, k
        k += 1
        assert k < 500
    monkeypatch.delenv("AFFT_TEST_FAIL_ALLOC")
    destroy(h)
    torch.cuda.synchronize()
    assert k > 2, "the creation should have made several allocations"
    assert torch.cuda.mem_get_info()[0] == free0
    return k - 1


@pytest.mark.parametrize("shape,dtype,axes", [
    ((3, 10007), 1, 0),      # multi-kernel Bluestein in FP64: tables, FP64 filter sub-plan, workspace
    ((4096, 4096), 0, 1),    # rows and a split column axis
    ((2, 1 << 22), 0, 0),    # split six-step with its row table
    ((64, 257), 0, 0),       # fused Bluestein
])
def test_plan_create_allocation_failures_leak_nothing(gpu, monkeypatch, shape, dtype, axes):
    n = _fail_each_allocation(monkeypatch, lambda: _create(shape, dtype, axes),
                              lambda h: L.lib().afft_plan_destroy(h))
    assert n >= 2


def test_ragged_create_allocation_failures_leak_nothing(gpu, monkeypatch):
    lengths = [64, 4099, 20480, 1000, 10007, 3, 0, 40960, 257]
    n = _fail_each_allocation(monkeypatch, lambda: _create_ragged(lengths, 0),
                              lambda h: L.lib().afft_ragged_plan_destroy(h))
    assert n >= 5
