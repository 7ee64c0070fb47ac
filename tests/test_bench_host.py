"""Host logic of bench.py (no GPU): the seeded serving request stream and the event
sampling stride of the timed runs."""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402


def test_serving_stream_is_seeded_and_rate_only_scales_arrivals():
    a = bench.serving_requests()
    b = bench.serving_requests()
    c = bench.serving_requests(rate_per_s=3000.0)
    assert a == b
    assert len(a) == 256
    # the same jobs in the same order; only the arrival times change (by the rate ratio)
    assert [j[:4] for j in a] == [j[:4] for j in c]
    assert np.allclose([j[4] for j in a], [2 * j[4] for j in c])
    assert all(a[i][4] <= a[i + 1][4] for i in range(255))
    kinds = [j[0] for j in a]
    assert set(kinds) == {0, 1, 2, 3}  # wide-Fib, LogUp, Poseidon, XOR AIR
    for kind, lo, hi in ((0, 14, 20), (1, 16, 18), (2, 12, 16), (3, 12, 18)):
        logs = [j[1] for j in a if j[0] == kind]
        assert min(logs) >= lo and max(logs) <= hi
    assert all(j[2] <= j[1] for j in a if j[0] == 3)  # XOR log_step <= log_size


def test_sample_stride_rotates_over_the_launches_of_a_run():
    for n in range(1, 80):
        k = bench._sample_every(n)
        assert k >= 1
        if k > 1:
            assert n % k != 0
            # over k runs every launch position of a run is sampled once
            seen = {(r * n + i) % n for r in range(k) for i in range(n) if (r * n + i) % k == 0}
            assert seen == set(range(n))
        assert n / k < 8  # a handful of bracketed launches per run (about four when n >= 8)
