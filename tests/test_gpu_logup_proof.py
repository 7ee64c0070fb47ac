"""The XOR-lookup AIR as a whole proof on the GPU (stwo_rocm_prove_logup_xor): byte-equal to the
oracle's proof for several PCS configurations, accepted by the verifier restatement, the same
alone and in the serving mix."""
import hashlib
import sys

import pytest

import oracle
import stark_verifier as sv
import stwo_rocm

pytestmark = pytest.mark.gpu

CASES = [
    (16, 7, dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)),
    (16, 99, dict(pow_bits=4, log_blowup_factor=2, log_last_layer_degree_bound=1, n_queries=20)),
    (17, 3, dict(pow_bits=0, log_blowup_factor=1, log_last_layer_degree_bound=2, n_queries=40)),
    (18, 2024, dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)),
    (16, 5, dict(pow_bits=26, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=70)),
    (19, 11, dict(pow_bits=8, log_blowup_factor=3, log_last_layer_degree_bound=0, n_queries=9)),
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("log_n,seed,cfg", CASES,
                         ids=["16-functional", "16-blowup2", "17-last2", "18-functional", "16-secure", "19-blowup3"])
def test_logup_xor_proof_equals_oracle_and_verifies(ctx, log_n, seed, cfg):
    _say("logup proof", log_n, seed, cfg)
    got = ctx.prove_logup_xor(log_n, seed, **cfg)
    again = ctx.prove_logup_xor(log_n, seed, **cfg)
    assert got == again
    exp = oracle.prove_logup_xor(log_n, seed, **cfg)
    assert hashlib.sha256(got).hexdigest() == hashlib.sha256(exp).hexdigest()
    sv.verify_proof(got, "xor_logup", {"log_n": log_n}, cfg)


def test_logup_xor_proof_alone_and_in_the_serving_mix(ctx):
    jobs = [(4, 16, 0, 31, 0.0), (0, 14, 20, 5, 0.1), (4, 17, 0, 32, 0.2), (2, 12, 0, 0, 0.3), (4, 16, 0, 33, 0.4),
            (3, 13, 2, 1, 0.5), (4, 18, 0, 34, 0.6)]
    mix, st = stwo_rocm.serve(ctx, jobs, n_slots=4)
    assert all(r["status"] == 0 for r in mix) and st["max_in_flight"] > 1
    for job, r in zip(jobs, mix):
        alone, _ = stwo_rocm.serve(ctx, [job[:4] + (0.0,)], n_slots=1)
        assert alone[0]["digest"] == r["digest"], job
        if job[0] == 4:
            assert r["digest"] == hashlib.blake2s(oracle.prove_logup_xor(job[1], job[3])).hexdigest(), job


def test_logup_xor_proof_rejects_bad_statements(ctx):
    with pytest.raises(stwo_rocm.RocmError):
        ctx.prove_logup_xor(15, 1)  # the 2^16-row table does not fit
    with pytest.raises(stwo_rocm.RocmError):
        ctx.prove_logup_xor(16, 1, log_blowup_factor=0)
    assert len(ctx.prove_logup_xor(16, 1)) > 0  # the context survives
