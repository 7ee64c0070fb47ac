"""Many small wide-Fibonacci proofs in single launches (stwo_rocm_prove_wide_fibonacci_batch).

Every proof of a batch must be byte-identical to the one-proof path's proof of the same
statement, trace seed and PCS configuration (itself pinned to the reference's golden proofs and
to the oracle), and a sample of them to the oracle directly.  The cases cover the batch's
branches: the fused interpolation + extension (blowup 1) and the two-launch form (blowup 2),
trees that start with wide parent launches (n + blowup = 17), a FRI with no inner layer, the
one-workgroup tail, the quotient split for more than 128 columns, proof of work from 0 to 26
bits, and the reference's own (1, i) trace against its golden proof."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
import stark_verifier as sv
import stwo_rocm

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# (log_n_rows, sequence_len, seeds, pcs) -- pcs: pow_bits, log_blowup_factor, last-layer bound, queries
CASES = [
    (6, 4, [1, 2, 3, 4, 5], dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)),
    (8, 12, list(range(100, 164)), dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0,
                                         n_queries=3)),
    (7, 6, [9, 10, 11], dict(pow_bits=0, log_blowup_factor=2, log_last_layer_degree_bound=1, n_queries=20)),
    (2, 3, [5, 6], dict(pow_bits=4, log_blowup_factor=1, log_last_layer_degree_bound=1, n_queries=4)),
    (16, 3, [21, 22], dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)),
    (9, 130, [31, 32, 33], dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=2, n_queries=8)),
    (11, 16, [41, 42], dict(pow_bits=26, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=70)),
]
IDS = ["6x4-5", "8x12-64", "7x6-blowup2", "2x3-no-inner", "16x3-wide-tree", "9x130-split", "11x16-secure"]


@pytest.fixture(scope="module")
def ctx():
    with stwo_rocm.Context(0) as c:
        yield c


@pytest.mark.parametrize("n,w,seeds,pcs", CASES, ids=IDS)
def test_batch_equals_one_proof_path(ctx, n, w, seeds, pcs):
    print("batch %dx%d K=%d %s" % (n, w, len(seeds), pcs), flush=True)
    proofs, st = ctx.prove_wide_fibonacci_batch(n, w, seeds=seeds, trace_mode=1, **pcs)
    assert len(proofs) == len(seeds) and st["slice_bytes"] > 0
    for k, s in enumerate(seeds):
        one, _ = ctx.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=s, stats=False, **pcs)
        assert proofs[k] == one, "proof %d (seed %d) differs from the one-proof path" % (k, s)
    # distinct seeds give distinct proofs
    assert len({hashlib.sha256(p).hexdigest() for p in proofs}) == len(seeds)
    # the first and last against the oracle and the verifier
    for k in {0, len(seeds) - 1}:
        if n <= 11 and pcs["pow_bits"] <= 12:
            assert proofs[k] == oracle.prove_wide_fibonacci(n, w, trace_mode=1, trace_seed=seeds[k], **pcs)
        sv.verify_proof(proofs[k], "wide_fibonacci", {"log_n_rows": n, "sequence_len": w}, pcs)


def test_batch_reference_trace_is_the_golden_proof(ctx):
    idx = json.load(open(os.path.join(GOLD, "wide_fibonacci_goldens.json")))
    e = idx["wide_fibonacci_log10_seq8_pow10"]
    golden = open(os.path.join(GOLD, e["file"]), "rb").read()
    proofs, _ = ctx.prove_wide_fibonacci_batch(10, 8, n_proofs=4, trace_mode=0, pow_bits=10, log_blowup_factor=1,
                                               log_last_layer_degree_bound=0, n_queries=3)
    assert all(p == golden for p in proofs)


def test_batch_runs_back_to_back(ctx):
    """Two batches of different shapes in a row on one context, then the first again: the
    completion counters and cached blocks of one batch leave the next unaffected."""
    pcs = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
    a1, _ = ctx.prove_wide_fibonacci_batch(9, 8, seeds=[1, 2, 3], **pcs)
    b, _ = ctx.prove_wide_fibonacci_batch(12, 5, seeds=[7] * 9, **pcs)
    a2, _ = ctx.prove_wide_fibonacci_batch(9, 8, seeds=[1, 2, 3], **pcs)
    assert a1 == a2
    assert len(set(b)) == 1
    one, _ = ctx.prove_wide_fibonacci(12, 5, trace_mode=1, trace_seed=7, stats=False, **pcs)
    assert b[0] == one


def test_batch_rejects_unsupported_shapes(ctx):
    pcs = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
    with pytest.raises(stwo_rocm.RocmError):
        ctx.prove_wide_fibonacci_batch(17, 8, seeds=[1], **pcs)  # n + blowup = 18
    with pytest.raises(stwo_rocm.RocmError):
        ctx.prove_wide_fibonacci_batch(12, 8, seeds=[1], pow_bits=10, log_blowup_factor=1,
                                       log_last_layer_degree_bound=10, n_queries=3)  # last layer 2^11
    with pytest.raises(stwo_rocm.RocmError):
        ctx.prove_wide_fibonacci_batch(8, 2, seeds=[1], **pcs)  # sequence_len < 3
    # the context stays usable
    p, _ = ctx.prove_wide_fibonacci_batch(8, 4, seeds=[3], **pcs)
    one, _ = ctx.prove_wide_fibonacci(8, 4, trace_mode=1, trace_seed=3, stats=False, **pcs)
    assert p[0] == one


def test_batch_plan_graph_replays_equal_eager(ctx):
    """A batch plan run eagerly, recorded into a hipGraph and replayed with new seeds, with the
    program's own eager work and synchronisations between the replays: every replay's proofs
    equal the one-proof path's for its seeds (the graph carries no memset node)."""
    pcs = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
    with ctx.batch_prover(8, 6, 16, **pcs) as bp:
        for rep, seeds in enumerate([list(range(1, 17)), list(range(100, 116)), list(range(1, 17)),
                                     [7] * 16]):
            print("plan run", rep, flush=True)
            got, st = bp.run(seeds, graph=True)
            eager, _ = bp.run(seeds, graph=False)
            assert got == eager
            one, _ = ctx.prove_wide_fibonacci(8, 6, trace_mode=1, trace_seed=seeds[3], stats=False, **pcs)
            assert got[3] == one
            ctx.synchronize()
        other, _ = ctx.prove_wide_fibonacci_batch(8, 6, seeds=list(range(100, 116)), **pcs)
        again, _ = bp.run(list(range(100, 116)), graph=True)
        assert again == other


def test_batch_plan_replay_after_the_tower_moves():
    """A recorded plan reads the context's twiddle tower.  Growing the tower (a larger
    transform on the context) returns the old arrays to the arena cache and release_cached
    frees them; the next graph run must re-record against the new tower, not read freed
    memory, and its proofs stay equal to the first run's."""
    pcs = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
    seeds = list(range(50, 58))
    with stwo_rocm.Context(0) as c2:  # a fresh context: its tower covers only this plan
        with c2.batch_prover(9, 5, 8, **pcs) as bp:
            first, _ = bp.run(seeds, graph=True)
            builds = c2.runtime_info()["tower_builds"]
            c2.prepare_twiddles(24)  # the tower grows: the old arrays go to the cache
            c2.release_cached()      # ... and are freed
            assert c2.runtime_info()["tower_builds"] == builds + 1
            junk = [c2.to_device(np.full(1 << 20, 0x5A5A5A5A, dtype=np.uint32)) for _ in range(8)]
            try:
                for _ in range(2):
                    again, _ = bp.run(seeds, graph=True)
                    assert again == first
            finally:
                for j in junk:
                    j.free()
        one, _ = c2.prove_wide_fibonacci(9, 5, trace_mode=1, trace_seed=seeds[2], stats=False, **pcs)
        assert first[2] == one


def test_batch_plan_without_graph_form(ctx):
    """More than 120 columns on at most 2^16 rows: the quotients use a temporary buffer, so the
    plan runs eagerly even when a graph is asked for; results are unchanged."""
    pcs = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
    with ctx.batch_prover(7, 125, 3, **pcs) as bp:
        a, _ = bp.run([4, 5, 6], graph=True)
        b, _ = bp.run([4, 5, 6], graph=True)
    assert a == b
    one, _ = ctx.prove_wide_fibonacci(7, 125, trace_mode=1, trace_seed=5, stats=False, **pcs)
    assert a[1] == one
