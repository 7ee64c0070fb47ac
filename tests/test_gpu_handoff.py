"""Long runs of the last-workgroup hand-offs.

The multi-workgroup kernels exchange per-workgroup results through device-coherent stores and
a completion counter whose last incrementer finishes the work: the tree-layer launches of the
FRI cascade (subtree roots, then the channel step) and the GKR rounds above the one-workgroup
size (per-workgroup sums).  These tests run the same cascade and the same GKR batch proof
hundreds of times in a row, with launches of other sizes in between so that the workgroup
schedule varies, and require every run to equal the restatement / the oracle."""
import sys

import numpy as np
import pytest

import oracle
import stark_verifier as sv
from test_gpu_hooks import _commit_layers_restated, _m31
from test_gpu_logup import _proof_case, dev

pytestmark = pytest.mark.gpu

CASCADE_RUNS = 2000
GKR_RUNS = 300


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cascade(ctx, seed, d_line, log_size, last_log):
    ch = ctx.channel(seed, 0)
    below = [ctx.alloc(16 << (log_size - 1 - k)) for k in range(log_size - last_log - 1)]
    last = ctx.alloc(16 << last_log)
    trees = ctx.commit_fri_layers(ch, d_line, log_size, last_log, below, last)
    roots = [t.root for t in trees]
    state = ch.state()
    for x in trees + below + [last]:
        x.free()
    ch.close()
    return roots, state


def test_fri_cascade_long_run(ctx):
    """2^16 line down to 2^1: layers 2^15..2^11 are one tree-layer launch each over many
    workgroups, the rest the one-workgroup tail; a 2^13 cascade runs between the runs."""
    rng = np.random.default_rng(4242)
    big, small = (16, 1), (13, 2)
    line_b, line_s = _m31(rng, 4 << big[0]), _m31(rng, 4 << small[0])
    seed_b, seed_s = rng.bytes(32), rng.bytes(32)
    ref = sv.Channel()
    ref.digest, ref.n_draws = seed_b, 0
    _, roots_b = _commit_layers_restated(line_b, big[0], big[1], ref)
    state_b = (ref.digest, ref.n_draws)
    ref = sv.Channel()
    ref.digest, ref.n_draws = seed_s, 0
    _, roots_s = _commit_layers_restated(line_s, small[0], small[1], ref)
    d_b, d_s = ctx.to_device(line_b), ctx.to_device(line_s)
    try:
        for run in range(CASCADE_RUNS):
            roots, state = _cascade(ctx, seed_b, d_b, *big)
            assert roots == roots_b and state == state_b, "cascade run %d differs" % run
            if run % 3 == 0:
                roots, _ = _cascade(ctx, seed_s, d_s, *small)
                assert roots == roots_s, "interleaved cascade after run %d differs" % run
            if run % 250 == 0:
                _say("cascade run", run)
    finally:
        d_b.free()
        d_s.free()


def test_gkr_batch_long_run(ctx):
    """A batch whose first layers run over many workgroups (2^10 and 2^9 values), proven again
    and again with a one-workgroup batch in between."""
    insts = _proof_case(ctx, "mixed_big")
    small = _proof_case(ctx, "mixed")
    digest = bytes(range(32, 64))
    ref = oracle.gkr_prove(insts, digest)
    ref_small = oracle.gkr_prove(small, digest)
    keep = []

    def on_device(cases):
        out = []
        for kind, num, den, nv in cases:
            dn = dev(ctx, num) if num is not None else None
            dd = dev(ctx, den)
            keep.extend(x for x in (dn, dd) if x is not None)
            out.append((kind, dn, dd, nv))
        return out

    d_insts, d_small = on_device(insts), on_device(small)
    try:
        for run in range(GKR_RUNS):
            words = ctx.gkr_prove_batch(d_insts, digest)
            assert np.array_equal(words, ref), "GKR run %d differs" % run
            if run % 2 == 0:
                assert np.array_equal(ctx.gkr_prove_batch(d_small, digest), ref_small)
            if run % 50 == 0:
                _say("gkr run", run)
    finally:
        for x in keep:
            x.free()
