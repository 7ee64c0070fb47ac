"""Poseidon AIR proofs (src/examples/poseidon.zig) on the GPU through the C-ABI.

- Byte equality with both committed golden proofs of the reference CPU backend
  (log_n_instances 10 and 13, functional protocol).
- Byte equality with the oracle on configurations no golden covers: the serving sizes
  (log_n_instances 12..16), the smallest trace (2 rows), other blowups, last-layer bounds,
  query counts and PoW difficulties.
- Rejected statements leave the context usable.
"""
import hashlib

import pytest

import oracle
import stwo_rocm
from conftest import load_poseidon_goldens, poseidon_kwargs

pytestmark = pytest.mark.gpu
GOLDENS = load_poseidon_goldens()


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


def test_poseidon_invalid_statement_and_context_survives(ctx):
    with pytest.raises(stwo_rocm.RocmError) as e:
        ctx.prove_poseidon(3)  # a trace of 2^0 rows is below the FRI configuration
    assert e.value.code == 1
    g = GOLDENS[0]
    proof, _ = ctx.prove_poseidon(**poseidon_kwargs(g))
    assert proof == g["proof"]
