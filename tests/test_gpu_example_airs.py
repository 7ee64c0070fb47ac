"""Poseidon AIR (src/examples/poseidon.zig) and periodic XOR AIR (src/examples/xor.zig)
proofs on the GPU through the C-ABI.

- Byte equality with the committed golden proofs of the reference CPU backend
  (Poseidon log_n_instances 10 and 13; XOR log_size 14 and 16 with step 4, offset 3;
  functional protocol).
- Byte equality with the oracle on configurations no golden covers: the serving sizes,
  the smallest traces, other steps and offsets (including offsets above p and a step of
  the whole trace), blowups, last-layer bounds, query counts and PoW difficulties; each
  such proof is also accepted by the verifier restatement (tests/stark_verifier.py).
- Rejected statements leave the context usable.
"""
import hashlib

import pytest

import oracle
import stark_verifier as sv
import stwo_rocm
from conftest import load_poseidon_goldens, load_xor_goldens, poseidon_kwargs, xor_kwargs

pytestmark = pytest.mark.gpu
GOLDENS = load_poseidon_goldens()


def _pcs(case):
    return dict(pow_bits=case.get("pow_bits", 10), log_blowup_factor=case.get("log_blowup_factor", 1),
                log_last_layer_degree_bound=case.get("log_last_layer_degree_bound", 0),
                n_queries=case.get("n_queries", 3))
XOR_GOLDENS = load_xor_goldens()


@pytest.mark.parametrize("g", GOLDENS, ids=[g["name"] for g in GOLDENS])
def test_gpu_poseidon_proof_equals_golden(ctx, g):
    print("golden", g["name"], flush=True)
    proof, stats = ctx.prove_poseidon(**poseidon_kwargs(g))
    assert hashlib.sha256(proof).hexdigest() == g["sha256"]
    assert proof == g["proof"]
    print("  gpu ms %.2f host ms %.2f" % (stats["total_gpu_ms"], stats["total_host_ms"]), flush=True)


CASES = [
    dict(log_n_instances=12),
    dict(log_n_instances=16),
    dict(log_n_instances=4, pow_bits=2, n_queries=2),
    dict(log_n_instances=11, pow_bits=4, log_blowup_factor=2, n_queries=7),
    dict(log_n_instances=14, pow_bits=12, log_last_layer_degree_bound=2, n_queries=20),
    dict(log_n_instances=9, pow_bits=1, log_blowup_factor=3, log_last_layer_degree_bound=1, n_queries=5),
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_gpu_poseidon_proof_equals_oracle(ctx, case):
    print("case", case, flush=True)
    exp = oracle.prove_poseidon(**case)
    got, _ = ctx.prove_poseidon(stats=False, **case)
    assert got == exp
    sv.verify_proof(got, "poseidon", dict(log_n_instances=case["log_n_instances"]), _pcs(case))


def test_poseidon_invalid_statement_and_context_survives(ctx):
    with pytest.raises(stwo_rocm.RocmError) as e:
        ctx.prove_poseidon(3)  # a trace of 2^0 rows is below the FRI configuration
    assert e.value.code == 1
    g = GOLDENS[0]
    proof, _ = ctx.prove_poseidon(**poseidon_kwargs(g))
    assert proof == g["proof"]


@pytest.mark.parametrize("g", XOR_GOLDENS, ids=[g["name"] for g in XOR_GOLDENS])
def test_gpu_xor_proof_equals_golden(ctx, g):
    print("golden", g["name"], flush=True)
    proof, _ = ctx.prove_xor(**xor_kwargs(g))
    assert hashlib.sha256(proof).hexdigest() == g["sha256"]
    assert proof == g["proof"]


def test_example_air_proofs_fit_a_small_hbm_budget():
    # three columns of 2^16 rows and 1264 columns of 2^10 rows: each proof's device footprint
    # stays far below 256 MiB (a constraint-count slip once sized a per-constraint buffer at
    # 16 * (2^32 - 1) bytes, which a large card could still satisfy)
    small = stwo_rocm.Context(0, hbm_budget_bytes=256 << 20)
    try:
        g = XOR_GOLDENS[-1]
        assert small.prove_xor(**xor_kwargs(g))[0] == g["proof"]
        g = GOLDENS[-1]
        assert small.prove_poseidon(**poseidon_kwargs(g))[0] == g["proof"]
    finally:
        small.close()


XOR_CASES = [
    dict(log_size=12, log_step=2, offset=3),
    dict(log_size=18, log_step=5, offset=17),
    dict(log_size=2, log_step=0, offset=0, pow_bits=2, n_queries=2),
    dict(log_size=10, log_step=10, offset=(1 << 40) + 5, pow_bits=3),
    dict(log_size=13, log_step=3, offset=2, log_blowup_factor=2, log_last_layer_degree_bound=1, n_queries=9),
    dict(log_size=9, log_step=1, offset=1, pow_bits=12, log_blowup_factor=3, log_last_layer_degree_bound=2,
         n_queries=20),
]


@pytest.mark.parametrize("case", XOR_CASES, ids=[str(i) for i in range(len(XOR_CASES))])
def test_gpu_xor_proof_equals_oracle(ctx, case):
    print("xor case", case, flush=True)
    exp = oracle.prove_xor(**case)
    got, _ = ctx.prove_xor(stats=False, **case)
    assert got == exp
    sv.verify_proof(got, "xor", dict(log_size=case["log_size"], log_step=case["log_step"], offset=case["offset"]),
                    _pcs(case))


def test_xor_invalid_statement_and_context_survives(ctx):
    with pytest.raises(stwo_rocm.RocmError) as e:
        ctx.prove_xor(8, 9, 0)  # log_step above log_size
    assert e.value.code == 1
    g = XOR_GOLDENS[0]
    proof, _ = ctx.prove_xor(**xor_kwargs(g))
    assert proof == g["proof"]
