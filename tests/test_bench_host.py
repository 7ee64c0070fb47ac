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
    assert set(kinds) == {0, 2, 3, 4}  # wide-Fib, Poseidon, XOR AIR, XOR-lookup AIR
    for kind, lo, hi in ((0, 14, 20), (4, 16, 18), (2, 12, 16), (3, 12, 18)):
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


def test_sample_stride_for_one_bracketed_launch_per_proof():
    # the LDE unit has 4 launches per proof; about one of them carries events per proof
    k = bench._sample_every(4, target=1)
    assert k > 1 and 4 % k != 0 and k % 4 != 0


def test_batched_column_logs_follow_the_seeded_rule():
    # SplitMix64 of state 0 (Vigna's splitmix64.c) pins the mixer; logs are in range and seeded
    assert bench._splitmix64(0) == 0xE220A8397B1DCDAF
    logs = bench.batched_column_logs(bench.BATCHED_SPEC)
    assert len(logs) == 64 and min(logs) >= 14 and max(logs) <= 22 and len(set(logs)) > 4
    assert logs == bench.batched_column_logs(bench.BATCHED_SPEC)
    M = (1 << 64) - 1
    assert logs[5] == 14 + bench._splitmix64(1 ^ 0xA5A5A5A5 ^ ((5 * 0x9E3779B97F4A7C15) & M)) % 9


def test_lde_unit_classes_and_unit_figures():
    prof = {"fft_ifft_pass m=20 R=13 C=1 cols=100": {}, "fft_ifft_pass m=20 R=7 C=64 cols=100": {},
            "fft_lde_pass m=21 R=8 C=32 cols=100": {}, "fft_lde_pass m=21 R=13 C=1 cols=100": {},
            "fft_lde_pass m=21 R=13 C=1 cols=8": {}, "fft_ifft_pass m=21 R=13 C=1 cols=4": {}, "quotients": {}}
    cls = bench.lde_unit_classes(prof, 20, 100, 1)
    assert len(cls) == 4 and all(k.endswith("cols=100") for k in cls)
    ms = {k: 0.25 for k in cls}
    counters = {"traffic": {k: 1.0e9 for k in cls}, "valu": {k: 1.0e8 for k in cls}}
    u = bench.unit_figures("x", cls, 1.6e9, ms, counters)
    assert abs(u["achieved"] - 1600.0) < 1e-6 and abs(u["frac"] - 0.2) < 1e-9
    assert u["traffic"] == 4e9 and abs(u["traffic_ratio"] - 2.5) < 1e-9
    assert abs(u["valu"]["bound_ms"] - 4e8 / (256 * 2.4e9) * 1e3) < 1e-3
    half = bench.unit_figures("x", cls[:1], 1.0e9, ms, counters, share={cls[0]: 0.5})
    assert abs(half["device_ms"] - 0.125) < 1e-9 and half["traffic"] == 5e8


def test_profiled_counters_parse_the_committed_summaries():
    c = bench.profiled_counters()
    assert set(c) == {"traffic", "valu"}
    assert all(v > 0 for v in c["traffic"].values()) and all(v > 0 for v in c["valu"].values())


def test_serving_oracle_subset_covers_every_kind():
    import bench

    jobs = bench.serving_requests()
    sub = bench.serving_oracle_subset(jobs)
    assert {jobs[i][0] for i in sub} == {j[0] for j in jobs}
    assert all(0 <= i < len(jobs) for i in sub) and sub == sorted(set(sub))
