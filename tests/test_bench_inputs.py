"""bench.py draws its inputs from the same counter-based SplitMix64 stream as the oracle's
generator (orc_gen_uniform / orc_gen_normal), so the benchmark's seeded inputs are the tests'."""
import os
import sys

import numpy as np

import oracle as O

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_uniform_is_the_oracle_stream_bit_for_bit():
    for seed in (0, 2, 4):
        for shape in ((7,), (3, 1000), (1, 5000)):
            np.testing.assert_array_equal(bench.splitmix_uniform(seed, shape, "complex128"), O.uniform(seed, shape))
    # across the chunk boundary of the vectorised generator
    n = (1 << 21) + 3
    np.testing.assert_array_equal(bench.splitmix_uniform(2, (n,), "complex64"), O.uniform(2, (n,), np.complex64))


def test_normal_is_the_oracle_stream_to_rounding():
    # Box-Muller through numpy's log/sqrt/cos/sin instead of the C library's: equal to a few ulp
    for seed in (3, 5, 6, 7):
        a = bench.splitmix_normal(seed, (4, 999), "complex128")
        b = O.normal(seed, (4, 999))
        assert np.max(np.abs(a - b)) <= 8 * np.finfo(np.float64).eps * max(1.0, np.max(np.abs(b)))
    n = (1 << 21) + 5
    a = bench.splitmix_normal(5, (n,), "complex64")
    b = O.normal(5, (n,), np.complex64)
    assert np.max(np.abs(a - b)) <= 2 * np.finfo(np.float32).eps * np.max(np.abs(b))


def test_ragged_lengths_follow_the_size_set():
    lengths = bench.ragged_lengths()
    assert len(lengths) == 10000
    assert set(lengths) <= set(bench.RAGGED_SIZES)
    assert 4e7 < sum(lengths) < 7e7
