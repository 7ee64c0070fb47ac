"""The benchmark configurations at their full sizes on the GPU, each against the oracle
(whose proofs the Python verifier restatement accepts) or, where a result is too large to
recompute, through size-independent properties.

A  wide-Fibonacci 2^12 x 16 under the secure protocol (PoW 26, 70 queries), seeded-uniform
   and reference-seed traces: byte-equal to the oracle, accepted by the verifier.
B  wide-Fibonacci 2^20 x 100, seeded-uniform pairs (the target run), functional and secure
   protocols: byte-equal to the oracle, accepted by the verifier.
C  batched primitive pipeline, 64 columns of log 14..22, blowup 4, 3 folds of a log-22 line:
   graph replay, eager launches and the autotuned shape all equal the oracle.
E  the 256-request serving stream at its size ranges (wide-Fibonacci 14..20 x 16..100, XOR AIR
   12..18, Poseidon 12..16, XOR-lookup AIR 16..18): every result is the same alone and in the
   mix, and a seeded subset equals the oracle.
"""
import hashlib
import sys

import numpy as np
import pytest

import bench
import oracle
import stark_verifier as sv
import stwo_rocm

pytestmark = pytest.mark.gpu
SECURE = stwo_rocm.PROTOCOLS["secure"]
FUNCTIONAL = stwo_rocm.PROTOCOLS["functional"]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("trace", [(1, 7), (1, 2026), (0, 0)], ids=["uniform7", "uniform2026", "reference_seeds"])
def test_config_a_secure_widefib_2p12x16(ctx, trace):
    mode, seed = trace
    _say("config A", trace)
    got, _ = ctx.prove_wide_fibonacci(12, 16, trace_mode=mode, trace_seed=seed, **SECURE)
    exp = oracle.prove_wide_fibonacci(12, 16, trace_mode=mode, trace_seed=seed, **SECURE)
    assert got == exp
    sv.verify_proof(got, "wide_fibonacci", dict(log_n_rows=12, sequence_len=16), SECURE)


@pytest.mark.parametrize("proto", ["functional", "secure"])
def test_config_b_seeded_uniform_widefib_2p20x100(ctx, proto):
    cfg = stwo_rocm.PROTOCOLS[proto]
    _say("config B", proto)
    got, _ = ctx.prove_wide_fibonacci(20, 100, trace_mode=1, trace_seed=20, **cfg)
    again, _ = ctx.prove_wide_fibonacci(20, 100, trace_mode=1, trace_seed=20, stats=False, **cfg)
    assert got == again
    exp = oracle.prove_wide_fibonacci(20, 100, trace_mode=1, trace_seed=20, **cfg)
    assert hashlib.sha256(got).hexdigest() == hashlib.sha256(exp).hexdigest()
    sv.verify_proof(got, "wide_fibonacci", dict(log_n_rows=20, sequence_len=100), cfg)


CONFIG_C = (64, 14, 22, 2, 22, 3, 1)


def test_config_c_batched_full_spec(ctx):
    _say("config C", CONFIG_C)
    ref = oracle.batched_pipeline(*CONFIG_C)
    bp = ctx.batched(*CONFIG_C)
    try:
        eager, _ = bp.run(1, graph=False)
        assert np.array_equal(eager, ref)
        words, _ = bp.run(2, graph=True)
        assert np.array_equal(words, ref)
        tile, levels, table = bp.autotune(reps=1)
        _say("  autotuned", tile, levels)
        words, _ = bp.run(1, graph=True)
        assert np.array_equal(words, ref)
        other, _ = bp.run(1, graph=False)
        assert np.array_equal(other, ref)
    finally:
        bp.set_tuning(13, 10)
        bp.close()


def _ref_digest(job):
    kind, n, w, seed, _ = job
    if kind == 0:
        return hashlib.blake2s(oracle.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seed)).hexdigest()
    if kind == 2:
        return hashlib.blake2s(oracle.prove_poseidon(n)).hexdigest()
    if kind == 3:
        return hashlib.blake2s(oracle.prove_xor(n, w, seed)).hexdigest()
    if kind == 4:
        return hashlib.blake2s(oracle.prove_logup_xor(n, seed)).hexdigest()
    return hashlib.blake2s(oracle.logup_xor_pipeline(n, seed).astype(np.uint32).tobytes()).hexdigest()


def test_config_e_serving_stream_alone_mix_oracle(ctx):
    jobs = bench.serving_requests()
    assert len(jobs) == 256
    kinds = {j[0] for j in jobs}
    assert kinds == {0, 2, 3, 4}
    wf = [j for j in jobs if j[0] == 0]
    assert min(j[1] for j in wf) == 14 and max(j[1] for j in wf) == 20
    _say("config E: mix of 256 requests over 16 slots")
    mix, st = stwo_rocm.serve(ctx, jobs, n_slots=16)
    assert st["n_slots"] == 16 and st["max_in_flight"] > 1
    assert all(r["status"] == 0 for r in mix)
    _say("  p50 %.2f ms p99 %.2f ms %.0f proofs/s" % (st["p50_ms"], st["p99_ms"], st["jobs_per_s"]))
    # every request alone, on one slot, with nothing else in flight
    for i, job in enumerate(jobs):
        alone, _ = stwo_rocm.serve(ctx, [job[:4] + (0.0,)], n_slots=1)
        assert alone[0]["digest"] == mix[i]["digest"], (i, job)
    # a seeded subset (every 16th request, one of each kind, and the largest proof) against the oracle
    subset = set(range(0, 256, 16))
    for k in (0, 2, 3, 4):
        subset.add(next(i for i, j in enumerate(jobs) if j[0] == k))
    subset.add(max((i for i, j in enumerate(jobs) if j[0] == 0), key=lambda i: (jobs[i][1], jobs[i][2])))
    for i in sorted(subset):
        _say("  oracle", i, jobs[i][:4])
        assert mix[i]["digest"] == _ref_digest(jobs[i]), (i, jobs[i])
