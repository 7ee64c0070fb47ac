"""GPU parity of each rocm primitive against the oracle, through the C-ABI.

Bit-exact equality for every integer output (M31 words, hashes, nonces).
Seeded inputs; sizes the oracle finishes in well under a second.
"""
import hashlib
import sys

import numpy as np
import pytest

import mathref as mr
import oracle

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_cols(rng, n_cols, log):
    return [rng.integers(0, mr.P, 1 << log, dtype=np.uint32) for _ in range(n_cols)]


def test_identity_is_gfx950(ctx):
    ident = ctx.identity()
    _say("device:", ident)
    assert "gfx950" in ident


@pytest.mark.parametrize("log", [3, 5, 11, 12, 16, 21])
def test_twiddle_tower_matches_oracle(ctx, log):
    tw, itw = ctx.read_twiddles(log)
    otw, oitw = oracle.twiddles(log)
    assert np.array_equal(tw, otw) and np.array_equal(itw, oitw)


@pytest.mark.parametrize("log", [1, 2, 3, 4, 7, 10, 11, 12, 13, 16, 20, 21, 22])
def test_interpolate_matches_oracle(ctx, log):
    rng = np.random.default_rng(100 + log)
    n_cols = 3 if log <= 20 else 2
    cols = rand_cols(rng, n_cols, log)
    bufs = [ctx.to_device(c) for c in cols]
    _say("interpolate log", log)
    ctx.interpolate_circle_buffers([b.ptr for b in bufs], log)
    for b, c in zip(bufs, cols):
        assert np.array_equal(b.download(), oracle.circle_interpolate(c, log))
        b.free()


@pytest.mark.parametrize("log,blowup", [(1, 1), (2, 1), (2, 2), (3, 1), (5, 2), (10, 1), (11, 1), (12, 2), (15, 1),
                                        (18, 2), (20, 1)])
def test_evaluate_lde_matches_oracle(ctx, log, blowup):
    rng = np.random.default_rng(200 + log * 4 + blowup)
    n_cols = 3
    cols = rand_cols(rng, n_cols, log)
    src = [ctx.to_device(c) for c in cols]
    dst = [ctx.alloc(4 << (log + blowup)) for _ in cols]
    _say("lde log", log, "blowup", blowup)
    ctx.evaluate_circle_buffers([b.ptr for b in src], log, blowup, [b.ptr for b in dst])
    for d, c in zip(dst, cols):
        assert np.array_equal(d.download(), oracle.circle_lde(c, log, blowup))
    for b in src + dst:
        b.free()


def test_merkle_mixed_sizes_matches_oracle(ctx):
    rng = np.random.default_rng(5)
    logs = [4, 1, 9, 4, 7, 9, 2, 9, 12, 3] + [12] * 20
    cols = [rng.integers(0, mr.P, 1 << k, dtype=np.uint32) for k in logs]
    bufs = [ctx.to_device(c) for c in cols]
    _say("merkle mixed", len(cols))
    tree = ctx.commit_merkle([b.ptr for b in bufs], logs)
    root, layers = oracle.merkle_commit(cols)
    assert tree.root == root
    assert tree.max_log_size() == 12
    idx = [0, 1, 77, 4095]
    assert b"".join(tree.read_hashes(12, idx)) == b"".join(layers[12][32 * i: 32 * i + 32] for i in idx)
    assert tree.read_hashes(5, [3])[0] == layers[5][96:128]
    tree.free()
    for b in bufs:
        b.free()


STAGED = {
    # (count, log) groups; every 16-column message block whose largest log grows starts a new
    # shared-prefix segment of the leaf hash
    "four_segments": [(16, 3), (16, 5), (16, 8), (5, 11)],
    "two_levels_tiny": [(17, 1), (1, 2)],
    "same_level_blocks": [(32, 6), (3, 7)],
    "full_last_block": [(16, 4), (16, 9), (16, 10)],
    "five_segments_wide": [(3, 8), (13, 9), (16, 10), (16, 12), (16, 14), (2, 16)],
    # segments of >= 8 message blocks on <= 2^14 leaves take the quad-lane leaf kernel: one that
    # starts from a previous segment's lifted state and is not the last one, and one that ends
    "quad_prefix_not_last": [(128, 6), (130, 9), (4, 12)],
    "quad_after_small_prefix": [(16, 3), (200, 10)],
}


@pytest.mark.parametrize("name", list(STAGED))
def test_merkle_staged_sizes_shared_prefixes(ctx, name):
    rng = np.random.default_rng(len(name))
    logs = [k for n, k in STAGED[name] for _ in range(n)]
    perm = rng.permutation(len(logs))  # commit in shuffled order: the library sorts stably
    logs = [logs[i] for i in perm]
    cols = [rng.integers(0, mr.P, 1 << k, dtype=np.uint32) for k in logs]
    bufs = [ctx.to_device(c) for c in cols]
    _say("merkle staged", name, len(cols))
    tree = ctx.commit_merkle([b.ptr for b in bufs], logs)
    root, layers = oracle.merkle_commit(cols)
    L = max(logs)
    assert tree.root == root
    idx = sorted({0, 1, 2, (1 << L) - 1, (1 << L) // 3})
    assert b"".join(tree.read_hashes(L, idx)) == b"".join(layers[L][32 * i: 32 * i + 32] for i in idx)
    # the two layers below the leaves held intermediate states before the parents were built
    for lay in (L - 1, L - 2):
        if lay >= 0:
            n = 1 << lay
            assert b"".join(tree.read_hashes(lay, list(range(min(n, 8))))) == layers[lay][: 32 * min(n, 8)]
    tree.free()
    for b in bufs:
        b.free()


@pytest.mark.parametrize("log", [9, 13, 18])
def test_merkle_every_level_matches_oracle(ctx, log):
    """Every level of the tree (the wide parent launches and the one-launch top reduction with
    its last-workgroup hand-off) at its first, last and two inner positions."""
    rng = np.random.default_rng(log)
    cols = rand_cols(rng, 3, log)
    bufs = [ctx.to_device(c) for c in cols]
    _say("merkle levels", log)
    tree = ctx.commit_merkle([b.ptr for b in bufs], [log] * 3)
    root, layers = oracle.merkle_commit(cols)
    assert tree.root == root
    for lay in range(log + 1):
        n = 1 << lay
        idx = sorted({0, n - 1, n // 3, (2 * n) // 3})
        assert b"".join(tree.read_hashes(lay, idx)) == b"".join(layers[lay][32 * i: 32 * i + 32] for i in idx), lay
    tree.free()
    for b in bufs:
        b.free()


@pytest.mark.parametrize("n_cols,log", [(0, 0), (1, 1), (1, 2), (3, 3), (5, 6), (16, 8), (17, 8), (2, 10), (2, 11),
                                        (100, 14), (8, 21)])
def test_merkle_uniform_matches_oracle(ctx, n_cols, log):
    rng = np.random.default_rng(n_cols * 31 + log)
    cols = rand_cols(rng, n_cols, log)
    bufs = [ctx.to_device(c) for c in cols]
    _say("merkle", n_cols, log)
    tree = ctx.commit_merkle([b.ptr for b in bufs], [log] * n_cols)
    assert tree.root == oracle.merkle_commit(cols)[0]
    tree.free()
    for b in bufs:
        b.free()


def test_merkle_decommit_witness_matches_oracle(ctx):
    rng = np.random.default_rng(6)
    cols = rand_cols(rng, 4, 10)
    bufs = [ctx.to_device(c) for c in cols]
    tree = ctx.commit_merkle([b.ptr for b in bufs], [10] * 4)
    _, layers = oracle.merkle_commit(cols)
    pos = [3, 4, 5, 600, 1023]
    wit = tree.decommit_witness(pos)
    # reference traversal (prover/vcs_lifted/decommit.zig:171-236) over oracle layers
    exp, prev = [], sorted(set(pos))
    for layer in range(9, -1, -1):
        cur, q = [], 0
        while q < len(prev):
            f = prev[q]
            sib = q + 1 < len(prev) and prev[q + 1] == f ^ 1
            if not sib:
                exp.append(layers[layer + 1][32 * (f ^ 1): 32 * (f ^ 1) + 32])
            cur.append(f >> 1)
            q += 2 if sib else 1
        prev = cur
    assert wit == exp
    tree.free()
    for b in bufs:
        b.free()


def test_merkle_parent_layer(ctx):
    rng = np.random.default_rng(7)
    children = rng.integers(0, 256, 64 * 8, dtype=np.uint8)
    d = ctx.to_device(children)
    out = ctx.alloc(32 * 8)
    ctx.merkle_parent_layer(d.ptr, 3, out.ptr)
    got = out.download(np.uint8)
    for i in range(8):
        exp = hashlib.blake2s(b"node" + bytes(60) + children[64 * i: 64 * i + 64].tobytes()).digest()
        assert got[32 * i: 32 * i + 32].tobytes() == exp


@pytest.mark.parametrize("log_src", [1, 2, 5, 12, 20])
def test_fold_line_matches_oracle(ctx, log_src):
    rng = np.random.default_rng(300 + log_src)
    src = rng.integers(0, mr.P, 4 << log_src, dtype=np.uint32)
    alpha = [int(v) for v in rng.integers(0, mr.P, 4)]
    d = ctx.to_device(src)
    out = ctx.alloc(16 << (log_src - 1))
    ctx.fold_line(d.ptr, log_src, alpha, out.ptr)
    exp = oracle.fold_line(src, log_src, 1 << (31 - log_src - 2), alpha)
    assert np.array_equal(out.download(), exp)
    d.free()
    out.free()


@pytest.mark.parametrize("log_src", [1, 3, 12, 21])
def test_fold_circle_into_line_matches_oracle(ctx, log_src):
    rng = np.random.default_rng(400 + log_src)
    coords = rand_cols(rng, 4, log_src)
    dst0 = rng.integers(0, mr.P, 4 << (log_src - 1), dtype=np.uint32)
    alpha = [int(v) for v in rng.integers(0, mr.P, 4)]
    dc = [ctx.to_device(c) for c in coords]
    dd = ctx.to_device(dst0)
    ctx.fold_circle_into_line([b.ptr for b in dc], log_src, alpha, dd.ptr)
    exp = oracle.fold_circle_into_line(coords, log_src, alpha, dst0)
    assert np.array_equal(dd.download(), exp)


@pytest.mark.parametrize("log,n_cols", [(0, 3), (1, 3), (5, 3), (14, 3), (20, 3), (12, 20), (16, 9)])
def test_eval_at_point_matches_oracle(ctx, log, n_cols):
    """Several column groups per chunk (the blocks of one chunk share an XCD) and chunk counts
    that are not a multiple of the XCD count."""
    rng = np.random.default_rng(500 + log)
    cols = rand_cols(rng, n_cols, log)
    pt = [int(v) for v in rng.integers(0, mr.P, 8)]
    bufs = [ctx.to_device(c) for c in cols]
    got = ctx.eval_at_point([b.ptr for b in bufs], log, pt)
    for i, c in enumerate(cols):
        assert list(got[i]) == oracle.eval_at_point(c, log, pt)


def test_batch_inverse_bitrev_accumulate(ctx):
    rng = np.random.default_rng(8)
    a = rng.integers(1, mr.P, 5000, dtype=np.uint32)
    d, o = ctx.to_device(a), ctx.alloc(4 * 5000)
    ctx.batch_inverse(d.ptr, 5000, o.ptr)
    inv = o.download()
    assert ((a.astype(np.uint64) * inv) % mr.P == 1).all()
    v = rng.integers(0, mr.P, 4 * 64, dtype=np.uint32)
    dv = ctx.to_device(v)
    ctx.bit_reverse(dv.ptr, 6, 4)
    br = dv.download().reshape(64, 4)
    vv = v.reshape(64, 4)
    assert all((br[mr.bitrev(i, 6)] == vv[i]).all() for i in range(64))
    x, y = rng.integers(0, mr.P, 1000, dtype=np.uint32), rng.integers(0, mr.P, 1000, dtype=np.uint32)
    dx, dy = ctx.to_device(x), ctx.to_device(y)
    ctx.accumulate(dx.ptr, dy.ptr, 1000)
    assert np.array_equal(dx.download(), ((x.astype(np.uint64) + y) % mr.P).astype(np.uint32))


@pytest.mark.parametrize("bits", [0, 1, 8, 14, 20])
def test_pow_grind_matches_oracle(ctx, bits):
    digest = hashlib.blake2s(b"grind%d" % bits).digest()
    assert ctx.grind(digest, bits) == oracle.pow_grind(digest, bits)


@pytest.mark.parametrize("log_n,seq,mode", [(5, 16, 0), (12, 100, 1), (20, 100, 0)])
def test_widefib_trace_matches_oracle(ctx, log_n, seq, mode):
    d = ctx.alloc(4 * seq << log_n)
    ctx.fill_quadratic_recurrence_trace(log_n, seq, d.ptr, mode=mode, seed=99)
    got = d.download().reshape(seq, 1 << log_n)
    assert np.array_equal(got, oracle.widefib_trace(log_n, seq, mode, 99))
    d.free()


def _quotients_ref(L, cols, logs, samples, alpha):
    """Row law of core/pcs/quotients/row_evaluation.zig:402-521 in Python (small L)."""
    # batches by point, first occurrence; powers in sample order
    batches, rp = [], (1, 0, 0, 0)
    for col, pt, val in samples:
        for b in batches:
            if b["pt"] == pt:
                break
        else:
            b = {"pt": pt, "e": [], "sa": (0, 0, 0, 0), "sb": (0, 0, 0, 0)}
            batches.append(b)
        conj = lambda q: (q[0], q[1], (-q[2]) % mr.P, (-q[3]) % mr.P)
        py = pt[4:]
        a = mr.qsub(conj(val), val)
        c = mr.qsub(conj(py), py)
        bb = mr.qsub(mr.qmul(val, c), mr.qmul(a, py))
        b["e"].append((col, mr.qmul(rp, c)))
        b["sa"] = mr.qadd(b["sa"], mr.qmul(rp, a))
        b["sb"] = mr.qadd(b["sb"], mr.qmul(rp, bb))
        rp = mr.qmul(rp, alpha)
    out = []
    for row in range(1 << L):
        x, y = mr.point_at_index(mr.domain_index(L, mr.bitrev(row, L)))
        acc = (0, 0, 0, 0)
        for b in batches:
            num = (0, 0, 0, 0)
            for col, cf in b["e"]:
                sh = L - logs[col] + 1
                v = int(cols[col][((row >> sh) << 1) + (row & 1)])
                num = mr.qadd(num, tuple(w * v % mr.P for w in cf))
            num = mr.qsub(num, mr.qadd(tuple(w * y % mr.P for w in b["sa"]), b["sb"]))
            px, py = b["pt"][:4], b["pt"][4:]
            # den = (prx - x) * piy - (pry - y) * pix  in CM31
            den = mr.csub(mr.cmul(((px[0] - x) % mr.P, px[1]), (py[2], py[3])),
                          mr.cmul(((py[0] - y) % mr.P, py[1]), (px[2], px[3])))
            inv = mr.qinv((den[0], den[1], 0, 0))
            acc = mr.qadd(acc, mr.qmul(num, inv))
        out.append(acc)
    return out


def test_quotients_match_row_law(ctx):
    rng = np.random.default_rng(9)
    L = 6
    logs = [6, 4, 6, 5]
    cols = [rng.integers(0, mr.P, 1 << k, dtype=np.uint32) for k in logs]
    pts = [tuple(int(v) for v in rng.integers(0, mr.P, 8)) for _ in range(2)]
    samples = [(0, pts[0], tuple(int(v) for v in rng.integers(0, mr.P, 4))),
               (1, pts[1], tuple(int(v) for v in rng.integers(0, mr.P, 4))),
               (2, pts[0], tuple(int(v) for v in rng.integers(0, mr.P, 4))),
               (3, pts[1], tuple(int(v) for v in rng.integers(0, mr.P, 4)))]
    alpha = tuple(int(v) for v in rng.integers(0, mr.P, 4))
    bufs = [ctx.to_device(c) for c in cols]
    outs = [ctx.alloc(4 << L) for _ in range(4)]
    tree = ctx.compute_quotients(L, [b.ptr for b in bufs], logs, [s[0] for s in samples],
                                 [list(s[1]) for s in samples], [list(s[2]) for s in samples], list(alpha),
                                 [o.ptr for o in outs], commit=True)
    got = np.stack([o.download() for o in outs], axis=1)
    exp = np.array(_quotients_ref(L, cols, logs, samples, alpha), dtype=np.uint32)
    assert np.array_equal(got, exp)
    assert tree.root == oracle.merkle_commit([got[:, k].copy() for k in range(4)])[0]
    tree.free()


def test_composition_matches_oracle_proof_path(ctx):
    """Composition on CanonicCoset(n+1) from the oracle's own LDE (B=1 path)."""
    rng = np.random.default_rng(10)
    n, W = 6, 8
    trace = oracle.widefib_trace(n, W, 1, 5)
    lde = [oracle.circle_lde(oracle.circle_interpolate(trace[c], n), n, 1) for c in range(W)]
    alpha = tuple(int(v) for v in rng.integers(0, mr.P, 4))
    bufs = [ctx.to_device(c) for c in lde]
    outs = [ctx.alloc(4 << (n + 1)) for _ in range(4)]
    ctx.compute_composition_evaluation([b.ptr for b in bufs], n, list(alpha), [o.ptr for o in outs])
    got = np.stack([o.download() for o in outs], axis=1)
    # reference law (component.zig:146-228) in Python
    N = W - 2
    pw = [(1, 0, 0, 0)]
    for _ in range(N - 1):
        pw.append(mr.qmul(pw[-1], alpha))

    def vanish(r):
        x, y = mr.point_at_index(mr.domain_index(n + 1, r))
        # p - coset.initial + coset.half_step with odds(n): initial = half_step = index 2^(31-n-1)
        g = mr.point_at_index(1 << (31 - n - 1))
        p = mr.padd(mr.padd((x, y), (g[0], (-g[1]) % mr.P)), g)
        xx = p[0]
        for _ in range(1, n):
            xx = mr.double_x(xx)
        return pow(xx, mr.P - 2, mr.P)

    den = [vanish(0), vanish(1)]
    for row in range(1 << (n + 1)):
        a, b = int(lde[0][row]), int(lde[1][row])
        acc = (0, 0, 0, 0)
        for i in range(N):
            cv = int(lde[i + 2][row])
            cons = (cv - a * a - b * b) % mr.P
            acc = mr.qadd(acc, tuple(w * cons % mr.P for w in pw[N - 1 - i]))
            a, b = b, cv
        d = den[row >> n]
        assert tuple(int(v) for v in got[row]) == tuple(w * d % mr.P for w in acc)


def test_unaligned_column_pointers_are_rejected(ctx):
    """Column entries read 16-byte words: a pointer 4 bytes into a buffer is refused on the host,
    before any kernel is launched."""
    import stwo_rocm
    rng = np.random.default_rng(9)
    col = rng.integers(0, mr.P, 1 << 6, dtype=np.uint32)
    buf = ctx.to_device(np.concatenate([col, col]))
    with pytest.raises(stwo_rocm.RocmError) as e:
        ctx.eval_at_point([buf.ptr + 4], 6, [1, 2, 3, 4, 5, 6, 7, 8])
    assert e.value.code == 1 and "16-byte aligned" in str(e.value)
    # an aligned column inside the same buffer (its second copy, 64 words in) works
    got = ctx.eval_at_point([buf.ptr + 4 * 64], 6, [1, 2, 3, 4, 5, 6, 7, 8])
    assert list(got[0]) == oracle.eval_at_point(col, 6, [1, 2, 3, 4, 5, 6, 7, 8])
    buf.free()
