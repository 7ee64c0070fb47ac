"""The resident hook surface of the C-ABI, hook by hook against the oracle and the Python
restatements, and whole proofs assembled only from hook calls in the generic prover's order
(tests/hook_prover.py) against the reference's golden proofs and the monolithic entry point."""
import hashlib
import sys

import numpy as np
import pytest

import hook_prover
import oracle
import stark_verifier as sv
import stwo_rocm
from conftest import golden_kwargs, load_goldens

pytestmark = pytest.mark.gpu
P = stwo_rocm.P


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _m31(rng, n):
    return rng.integers(0, P, size=n, dtype=np.uint32)


def test_device_channel_matches_restatement(ctx):
    rng = np.random.default_rng(11)
    ch = ctx.channel()
    ref = sv.Channel()
    try:
        for step in range(40):
            op = step % 7
            if op == 0:
                w = [int(x) for x in rng.integers(0, 2**32, size=int(rng.integers(0, 40)), dtype=np.uint64)]
                ch.mix_u32s(w)
                ref.mix_u32s(w)
            elif op == 1:
                v = int(rng.integers(0, 2**63))
                ch.mix_u64(v)
                ref.mix_u64(v)
            elif op == 2:
                f = [[int(x) for x in _m31(rng, 4)] for _ in range(int(rng.integers(0, 30)))]
                ch.mix_felts(f)
                ref.mix_felts(f)
            elif op == 3:
                r = rng.bytes(32)
                ch.mix_root(r)
                ref.mix_root(r)
            elif op in (4, 5):
                assert tuple(ch.draw_secure_felt()) == ref.draw_secure_felt()
            else:
                k = int(rng.integers(1, 4))
                got = ch.draw_u32s(k)
                exp = [w for _ in range(k) for w in ref.draw_u32s()]
                assert got == exp
            assert ch.state() == (ref.digest, ref.n_draws), step
        ch.mix_u32s([])  # an empty mix rehashes the digest
        ref.mix_u32s([])
        assert ch.state() == (ref.digest, ref.n_draws)
        d = ctx.channel(ref.digest, 5)
        assert d.state() == (ref.digest, 5)
        d.close()
    finally:
        ch.close()


@pytest.mark.parametrize("logs,blowup", [([5, 5, 5], 1), ([3, 7, 4, 7, 2], 2), ([12] * 6, 1), ([1, 2], 3),
                                         ([16, 10, 16], 1)])
def test_prepare_and_commit_owned_and_polys(ctx, logs, blowup):
    rng = np.random.default_rng(sum(logs) * 7 + blowup)
    vals = [_m31(rng, 1 << lg) for lg in logs]
    bufs = [ctx.to_device(v) for v in vals]
    ext = [ctx.alloc(4 << (lg + blowup)) for lg in logs]
    tree = ctx.prepare_and_commit_owned([int(b) for b in bufs], logs, blowup, [int(e) for e in ext])
    coeffs = [oracle.circle_interpolate(v, lg) for v, lg in zip(vals, logs)]
    ldes = [oracle.circle_lde(c, lg, blowup) for c, lg in zip(coeffs, logs)]
    for b, c in zip(bufs, coeffs):
        assert np.array_equal(b.download(), c)  # coefficients retained in place
    for e, l_ in zip(ext, ldes):
        assert np.array_equal(e.download(), l_)
    root, _ = oracle.merkle_commit(ldes)
    assert tree.root == root
    # the same commitment from the coefficients (prepareAndCommitPolys)
    ext2 = [ctx.alloc(4 << (lg + blowup)) for lg in logs]
    tree2 = ctx.prepare_and_commit_polys([int(b) for b in bufs], logs, blowup, [int(e) for e in ext2])
    assert tree2.root == root
    for e, l_ in zip(ext2, ldes):
        assert np.array_equal(e.download(), l_)
    for x in [tree, tree2] + bufs + ext + ext2:
        x.free()


def test_prepare_and_commit_without_columns(ctx):
    tree = ctx.prepare_and_commit_owned([], [], 1, [])
    assert tree.root == hashlib.blake2s(b"leaf" + bytes(60)).digest()
    tree.free()


@pytest.mark.parametrize("log", [1, 4, 10, 17])
def test_secure_column_and_line_commit(ctx, log):
    rng = np.random.default_rng(log)
    line = _m31(rng, 4 << log)
    d = ctx.to_device(line)
    coords = [ctx.alloc(4 << log) for _ in range(4)]
    ctx.secure_column_from_line(d, 1 << log, [int(c) for c in coords])
    q = line.reshape(-1, 4)
    for k in range(4):
        assert np.array_equal(coords[k].download(), q[:, k])
    tree = ctx.commit_line(d, log)
    root, _ = oracle.merkle_commit([q[:, k] for k in range(4)])
    assert tree.root == root
    for x in [tree, d] + coords:
        x.free()


def _half_odds_init(log):
    return 1 << (31 - log - 2)


@pytest.mark.parametrize("log_src,n_folds", [(2, 1), (6, 1), (9, 3), (14, 2), (18, 1)])
def test_fold_line_and_commit_next(ctx, log_src, n_folds):
    rng = np.random.default_rng(log_src * 10 + n_folds)
    src = _m31(rng, 4 << log_src)
    alpha = [int(x) for x in _m31(rng, 4)]
    d_src = ctx.to_device(src)
    dst_log = log_src - n_folds
    d_dst = ctx.alloc(16 << dst_log)
    tree = ctx.fold_line_and_commit_next(d_src, log_src, alpha, n_folds, d_dst)
    cur, a = src, tuple(alpha)
    for f in range(n_folds):
        cur = oracle.fold_line(cur, log_src - f, _half_odds_init(log_src - f), list(a))
        a = sv.qmul(a, a)
    assert np.array_equal(d_dst.download(), cur)
    q = cur.reshape(-1, 4)
    root, _ = oracle.merkle_commit([q[:, k] for k in range(4)])
    assert tree.root == root
    for x in (tree, d_src, d_dst):
        x.free()


def _commit_layers_restated(line, log_size, last_log, ref):
    """commitInnerLayers with fold_step 1 (fri.zig:497-586) on the Python channel and the oracle."""
    lines, roots = [], []
    cur, lg = line, log_size
    while lg > last_log:
        q = cur.reshape(-1, 4)
        root, _ = oracle.merkle_commit([q[:, k] for k in range(4)])
        roots.append(root)
        ref.mix_root(root)
        a = ref.draw_secure_felt()
        cur = oracle.fold_line(cur, lg, _half_odds_init(lg), list(a))
        lines.append(cur)
        lg -= 1
    return lines, roots


@pytest.mark.parametrize("log_size,last_log", [(2, 1), (5, 1), (11, 2), (12, 1), (13, 3), (17, 1), (17, 12),
                                               (20, 2)])
def test_commit_fri_layers(ctx, log_size, last_log):
    _say("commit_fri_layers", log_size, last_log)
    rng = np.random.default_rng(log_size * 31 + last_log)
    line = _m31(rng, 4 << log_size)
    seed = rng.bytes(32)
    ch = ctx.channel(seed, 3)
    ref = sv.Channel()
    ref.digest, ref.n_draws = seed, 3
    d_line = ctx.to_device(line)
    below = [ctx.alloc(16 << (log_size - 1 - k)) for k in range(log_size - last_log - 1)]
    last = ctx.alloc(16 << last_log)
    trees = ctx.commit_fri_layers(ch, d_line, log_size, last_log, below, last)
    lines, roots = _commit_layers_restated(line, log_size, last_log, ref)
    assert [t.root for t in trees] == roots
    for b, exp in zip(below, lines[:-1]):
        assert np.array_equal(b.download(), exp)
    assert np.array_equal(last.download(), lines[-1])
    assert ch.state() == (ref.digest, ref.n_draws)
    # every layer's tree is resident: its decommitment checks out against the restated root
    for t, ln, lg in zip(trees, [line] + lines[:-1], range(log_size, last_log, -1)):
        pos = sorted(set(int(x) for x in rng.integers(0, 1 << lg, size=5)))
        q = ln.reshape(-1, 4)
        sv.merkle_verify(t.root, [lg] * 4, pos, [[int(q[p, k]) for p in pos] for k in range(4)],
                         t.decommit_witness(pos))
    for x in trees + below + [last, d_line, ch]:
        x.close() if hasattr(x, "close") else x.free()


def test_gather_values_lifted(ctx):
    rng = np.random.default_rng(5)
    logs = [3, 6, 6, 9]
    L = 9
    cols = [_m31(rng, 1 << lg) for lg in logs]
    bufs = [ctx.to_device(c) for c in cols]
    pos = [0, 1, 5, 200, 511, 77, 77]
    got = ctx.gather_values([int(b) for b in bufs], logs, L, pos)
    for i, (c, lg) in enumerate(zip(cols, logs)):
        sh = L - lg + 1
        for j, p in enumerate(pos):
            row = p if lg == L else ((p >> sh) << 1) + (p & 1)
            assert got[i, j, 0] == c[row]
    line = _m31(rng, 4 << 7)
    d = ctx.to_device(line)
    g4 = ctx.gather_values([int(d)], [7], 7, [3, 100], words_per_value=4)
    assert np.array_equal(g4[0, 0], line[12:16]) and np.array_equal(g4[0, 1], line[400:404])
    for b in bufs + [d]:
        b.free()


@pytest.mark.parametrize("e", load_goldens(), ids=lambda e: e["name"])
def test_hook_assembled_proof_equals_golden(ctx, e):
    """The whole proof from hook calls only, in the generic prover's order: byte-equal to the
    reference's golden proof (the 2^20 x 100 one is sha256 e6609d05...)."""
    kw = golden_kwargs(e)
    _say("hook proof", e["name"])
    got = hook_prover.prove_wide_fibonacci(ctx, **kw)
    assert hashlib.sha256(got).hexdigest() == e["sha256"]
    assert got == e["proof"]


@pytest.mark.parametrize("kw", [
    dict(log_n_rows=12, sequence_len=16, pow_bits=26, n_queries=70, trace_mode=1, trace_seed=7),
    dict(log_n_rows=9, sequence_len=12, log_blowup_factor=2, log_last_layer_degree_bound=1, trace_mode=1,
         trace_seed=3),
    dict(log_n_rows=8, sequence_len=5, log_blowup_factor=3, log_last_layer_degree_bound=2, n_queries=20),
    dict(log_n_rows=14, sequence_len=40, pow_bits=4, log_last_layer_degree_bound=2, n_queries=9, trace_mode=1,
         trace_seed=99),
    dict(log_n_rows=2, sequence_len=3, pow_bits=0, n_queries=5),
], ids=["secure12x16", "blowup2_last1", "blowup3_last2", "14x40_last2", "tiny"])
def test_hook_assembled_proof_equals_monolithic(ctx, kw):
    _say("hook proof", kw)
    got = hook_prover.prove_wide_fibonacci(ctx, **kw)
    mono, _ = ctx.prove_wide_fibonacci(stats=False, **kw)
    assert got == mono
    cfg = dict(pow_bits=kw.get("pow_bits", 10), log_blowup_factor=kw.get("log_blowup_factor", 1),
               log_last_layer_degree_bound=kw.get("log_last_layer_degree_bound", 0),
               n_queries=kw.get("n_queries", 3))
    sv.verify_proof(got, "wide_fibonacci", dict(log_n_rows=kw["log_n_rows"], sequence_len=kw["sequence_len"]), cfg)


@pytest.mark.parametrize("log", [8, 13, 14, 16, 17, 18, 19, 20, 21, 22])
def test_interpolate_and_extend_fused_and_plain(ctx, log):
    """prepareAndCommitOwned at blowup 2: the interpolation's top pass and the extension's top
    pass run as one launch where the pass plans line up (logs 17..20 and 22 with 2^13 tiles) and
    as two transforms elsewhere; coefficients, extension and root equal the oracle's for both."""
    rng = np.random.default_rng(900 + log)
    vals = [_m31(rng, 1 << log) for _ in range(3)]
    bufs = [ctx.to_device(v) for v in vals]
    ext = [ctx.alloc(4 << (log + 1)) for _ in vals]
    ctx.profile_enable(True)
    tree = ctx.prepare_and_commit_owned([int(b) for b in bufs], [log] * 3, 1, [int(e) for e in ext])
    classes = ctx.profile_read()
    ctx.profile_enable(False)
    _say("log", log, "fused" if any(k.startswith("fft_ilde_pass") for k in classes) else "plain")
    coeffs = [oracle.circle_interpolate(v, log) for v in vals]
    ldes = [oracle.circle_lde(c, log, 1) for c in coeffs]
    for b, c in zip(bufs, coeffs):
        assert np.array_equal(b.download(), c)
    for e, l_ in zip(ext, ldes):
        assert np.array_equal(e.download(), l_)
    assert tree.root == oracle.merkle_commit(ldes)[0]
    if log in (17, 18, 19, 20, 22):
        assert any(k.startswith("fft_ilde_pass m=%d " % log) for k in classes), sorted(classes)
    for x in [tree] + bufs + ext:
        x.free()
