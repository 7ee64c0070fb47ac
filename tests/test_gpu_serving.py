"""Continuous-batching service on one GPU: every job's result is the same alone and in the mix,
wide-Fibonacci proofs in the mix equal the oracle's proofs byte for byte (Blake2s digest of the
proof bytes), Poseidon and periodic XOR AIR proofs equal the oracle's, LogUp jobs equal the oracle's LogUp
workload words, and the HBM budget bounds the
number of jobs admitted together."""
import hashlib
import sys

import numpy as np
import pytest

import oracle
import stwo_rocm

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


JOBS = [  # (kind, log_n, seq_len, seed, arrival_ms)
    (0, 10, 16, 101, 0.0), (1, 16, 0, 7, 0.5), (0, 8, 8, 102, 1.0), (2, 12, 0, 0, 1.2), (0, 12, 32, 103, 1.5),
    (0, 9, 64, 104, 2.0), (1, 16, 0, 8, 3.0), (2, 14, 0, 0, 3.2), (3, 14, 2, 3, 3.3), (0, 11, 24, 105, 3.5),
    (3, 12, 4, 9, 4.0), (0, 10, 100, 106, 6.0),
]


def _ref_digest(job):
    kind, n, w, seed, _ = job
    if kind == 0:
        proof = oracle.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seed)
        return hashlib.blake2s(proof).hexdigest()
    if kind == 2:
        return hashlib.blake2s(oracle.prove_poseidon(n)).hexdigest()
    if kind == 3:
        return hashlib.blake2s(oracle.prove_xor(n, w, seed)).hexdigest()
    words = oracle.logup_xor_pipeline(n, seed)
    return hashlib.blake2s(words.astype(np.uint32).tobytes()).hexdigest()


def test_serving_mix_equals_alone_and_oracle(ctx):
    _say("serving mix", len(JOBS))
    mix, st = stwo_rocm.serve(ctx, JOBS, n_slots=3)
    assert st["n_slots"] == 3 and 1 <= st["max_in_flight"] <= 3
    for job, r in zip(JOBS, mix):
        alone, _ = stwo_rocm.serve(ctx, [job[:4] + (0.0,)], n_slots=1)
        assert alone[0]["digest"] == r["digest"], job
        assert r["digest"] == _ref_digest(job), job
        assert r["end_ms"] >= r["start_ms"] >= r["arrival_ms"] - 1e-6


def test_serving_budget_limits_concurrency(ctx):
    jobs = [(0, 12, 32, 200 + i, 0.0) for i in range(6)]
    per_job = 16 * (1 << 12) * 32 + 1024 * (1 << 12) + (64 << 20)  # the admission estimate
    res, st = stwo_rocm.serve(ctx, jobs, n_slots=4, hbm_budget_bytes=2 * per_job)
    assert st["max_in_flight"] <= 2 and st["peak_admitted_bytes"] <= 2 * per_job
    ref, _ = stwo_rocm.serve(ctx, jobs, n_slots=1)
    assert [r["digest"] for r in res] == [r["digest"] for r in ref]


def test_serving_rejects_job_over_budget(ctx):
    with pytest.raises(stwo_rocm.RocmError):
        stwo_rocm.serve(ctx, [(0, 12, 32, 1, 0.0)], n_slots=1, hbm_budget_bytes=1 << 20)


def test_serving_slots_persist_and_rebuild_after_release(ctx):
    # slot contexts stay resident between calls; release_cached frees them and the next call
    # rebuilds them, with the same results
    jobs = [(0, 11, 24, 301, 0.0), (2, 12, 0, 0, 0.0), (3, 13, 3, 5, 0.0), (0, 9, 40, 302, 0.0)]
    a, _ = stwo_rocm.serve(ctx, jobs, n_slots=4)
    b, _ = stwo_rocm.serve(ctx, jobs, n_slots=2)
    ctx.release_cached()
    c, _ = stwo_rocm.serve(ctx, jobs, n_slots=4)
    assert [r["digest"] for r in a] == [r["digest"] for r in b] == [r["digest"] for r in c]
    assert a[0]["digest"] == _ref_digest(jobs[0])


def test_serving_budget_bounds_the_device_memory_the_slots_hold(ctx):
    """The budget covers what the slots really hold (live and cached arena blocks of every slot,
    towers included), not only the admission estimates: the peak of the device memory held by
    the slots stays within the budget, while the same jobs give the same results as without one."""
    kinds = [(0, 16, 64, 400 + i) for i in range(8)] + [(2, 14, 0, 0), (1, 16, 0, 9), (3, 15, 2, 7), (0, 17, 20, 99)]
    jobs = [k + (0.1 * i,) for i, k in enumerate(kinds)]
    budget = 1 << 30
    _say("serving under a 1 GiB budget")
    res, st = stwo_rocm.serve(ctx, jobs, n_slots=4, hbm_budget_bytes=budget)
    assert all(r["status"] == 0 for r in res)
    assert 0 < st["peak_device_bytes"] <= budget, st
    assert st["peak_admitted_bytes"] <= budget
    ref, st1 = stwo_rocm.serve(ctx, jobs, n_slots=1)
    assert [r["digest"] for r in res] == [r["digest"] for r in ref]
    # a budget below what one slot must hold beside the job is refused, not overrun
    with pytest.raises(stwo_rocm.RocmError):
        stwo_rocm.serve(ctx, [(0, 16, 64, 1, 0.0)], n_slots=1, hbm_budget_bytes=64 << 20)
