"""GPU parity of the batched primitive pipeline (many uneven columns LDE'd and committed in one
lifted tree, then FRI fold_line rounds on a device channel) against the oracle, for the captured
hipGraph, the same launches issued eagerly, and every launch-shape tuning.

Graph replays are checked one by one with eager work between them (a replay must not depend on
state left by the previous one: the device channel is cleared by a kernel inside the graph).
"""
import sys

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


SPECS = {
    "small": (64, 8, 13, 2, 12, 3, 7),
    "mixed": (16, 12, 18, 2, 18, 3, 8),
    "blowup1": (9, 5, 15, 1, 10, 2, 9),
}


@pytest.mark.parametrize("name", list(SPECS))
def test_batched_graph_and_eager_match_oracle(ctx, name):
    spec = SPECS[name]
    _say("batched", name, spec)
    ref = oracle.batched_pipeline(*spec)
    bp = ctx.batched(*spec)
    try:
        eager, _ = bp.run(1, graph=False)
        assert np.array_equal(eager, ref)
        for _ in range(3):
            words, _ = bp.run(1, graph=True)
            assert np.array_equal(words, ref)
            other, _ = bp.run(1, graph=False)  # eager work between replays
            assert np.array_equal(other, ref)
        words, _ = bp.run(4, graph=True)
        assert np.array_equal(words, ref)
    finally:
        bp.close()


@pytest.mark.parametrize("tile,levels", [(11, 6), (12, 8), (13, 10), (11, 10), (13, 2)])
def test_batched_every_tuning_is_bit_identical(ctx, tile, levels):
    spec = SPECS["small"]
    _say("tuning", tile, levels)
    ref = oracle.batched_pipeline(*spec)
    bp = ctx.batched(*spec)
    try:
        bp.set_tuning(tile, levels)
        words, _ = bp.run(2, graph=True)
        assert np.array_equal(words, ref)
    finally:
        bp.set_tuning(13, 10)
        bp.close()


def test_batched_autotune_picks_a_candidate_and_keeps_results(ctx):
    spec = SPECS["mixed"]
    ref = oracle.batched_pipeline(*spec)
    bp = ctx.batched(*spec)
    try:
        tile, levels, table = bp.autotune(reps=2)
        assert len(table) == 9 and (tile, levels) in {(t["fft_tile_log"], t["merkle_levels"]) for t in table}
        words, ms = bp.run(2, graph=True)
        assert np.array_equal(words, ref) and ms > 0
    finally:
        bp.set_tuning(13, 10)
        bp.close()


def test_batched_replay_after_the_tower_moves_and_tuning_stays_in_the_pipeline():
    """A captured graph reads the context's twiddle tower.  Growing the tower (a larger
    transform on the context) returns the old arrays to the arena cache and release_cached
    frees them; the next replay must recapture against the new tower, not read freed memory.
    The pipeline's launch shapes must not leak into the context's other work."""
    import stwo_rocm

    spec = SPECS["small"]
    ref = oracle.batched_pipeline(*spec)
    with stwo_rocm.Context(0) as c2:  # a fresh context: its tower covers only this pipeline
        bp = c2.batched(*spec)
        try:
            bp.set_tuning(11, 6)
            words, _ = bp.run(1, graph=True)
            assert np.array_equal(words, ref)
            c2.prepare_twiddles(24)  # the tower grows: the old arrays go to the cache
            c2.release_cached()      # ... and are freed
            junk = [c2.to_device(np.full(1 << 20, 0x5A5A5A5A, dtype=np.uint32)) for _ in range(8)]
            for _ in range(2):
                words, _ = bp.run(1, graph=True)
                assert np.array_equal(words, ref)
            for j in junk:
                j.free()
            # the context kept its own shapes: a proof's wide Merkle levels run 10 per launch
            c2.profile_enable(True)
            c2.prove_wide_fibonacci(18, 8, stats=False)
            classes = c2.profile_read()
            c2.profile_enable(False)
            assert any(k.startswith("merkle_parents L=19 levels=10") for k in classes), sorted(classes)
            assert not any("levels=6" in k for k in classes)
        finally:
            bp.close()
