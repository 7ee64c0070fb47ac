"""GPU parity of the LogUp / GKR column operations against the oracle, through the C-ABI.

Bit-exact equality of every word: the XOR-lookup trace and multiplicities, the
relation denominators, the cumulative (prefix-sum) column and its claimed sum,
the GKR hooks (eq evals, layer halving, variable fixing, round polynomials) and
whole batch GKR proofs (word streams, including the final channel digest).
Sizes run from ragged and single-element inputs to the 2^18-row LogUp workload.
"""
import sys

import numpy as np
import pytest

import oracle
from stwo_rocm import RocmError

pytestmark = pytest.mark.gpu
P = oracle.P


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rand_q(rng, n):
    return rng.integers(0, P, size=4 * n, dtype=np.uint64).astype(np.uint32)


def rand_m(rng, n, hi=P):
    return rng.integers(0, hi, size=n, dtype=np.uint64).astype(np.uint32)


def dev(ctx, arr):
    return ctx.to_device(np.ascontiguousarray(arr, dtype=np.uint32))


@pytest.mark.parametrize("log_size", [1, 12, 18])
def test_xor_trace_matches_oracle(ctx, log_size):
    _say("xor trace", log_size)
    n = 1 << log_size
    bufs = [ctx.alloc(4 * n) for _ in range(3)]
    dm = ctx.alloc(4 << 16)
    ctx.xor_trace(log_size, 0xABCDEF, *bufs, dm)
    a, b, c, m = oracle.xor_trace(log_size, 0xABCDEF)
    for buf, ref in zip(bufs, (a, b, c)):
        assert np.array_equal(buf.download(count=n), ref)
    assert np.array_equal(dm.download(count=1 << 16), m)


@pytest.mark.parametrize("table", [False, True])
def test_logup_denominators_match_oracle(ctx, table):
    _say("denominators", table)
    rng = np.random.default_rng(2)
    z, alpha = rand_q(rng, 1), rand_q(rng, 1)
    if table:
        n = 1 << 16
        dd = ctx.alloc(16 * n)
        ctx.logup_denominators(None, None, None, n, z, alpha, dd)
        ref = oracle.logup_denominators(None, None, None, z, alpha)
    else:
        a, b, c, _ = oracle.xor_trace(10, 5)
        n = len(a)
        da, db, dc = dev(ctx, a), dev(ctx, b), dev(ctx, c)
        dd = ctx.alloc(16 * n)
        ctx.logup_denominators(da, db, dc, n, z, alpha, dd)
        ref = oracle.logup_denominators(a, b, c, z, alpha)
    assert np.array_equal(dd.download(count=4 * n), ref)


@pytest.mark.parametrize("num_kind,n", [(0, 1), (1, 100), (2, 2048), (0, 2049), (1, 5000), (2, 1 << 18),
                                        (0, 1 << 20)])
def test_cumulative_column_matches_oracle(ctx, num_kind, n):
    _say("cumulative", num_kind, n)
    rng = np.random.default_rng(n + num_kind)
    den = rand_q(rng, n)
    num = rand_q(rng, n) if num_kind == 0 else rand_m(rng, n) if num_kind == 1 else None
    dd = dev(ctx, den)
    dn = dev(ctx, num) if num is not None else None
    ds = ctx.alloc(16 * n)
    claimed = ctx.logup_cumulative(num_kind, dn, dd, n, ds)
    ref_sums, ref_claimed = oracle.logup_cumulative(num_kind, num, den)
    assert np.array_equal(ds.download(count=4 * n), ref_sums)
    assert np.array_equal(claimed, ref_claimed)


def test_cumulative_column_rejects_zero_denominator(ctx):
    rng = np.random.default_rng(9)
    den = rand_q(rng, 300)
    den[4 * 123:4 * 124] = 0
    dd = dev(ctx, den)
    ds = ctx.alloc(16 * 300)
    with pytest.raises(RocmError):
        ctx.logup_cumulative(2, None, dd, 300, ds)


@pytest.mark.parametrize("n", [0, 1, 5, 17])
def test_gen_eq_evals_matches_oracle(ctx, n):
    _say("eq evals", n)
    rng = np.random.default_rng(40 + n)
    y, scale = rand_q(rng, n), rand_q(rng, 1)
    out = ctx.alloc(16 << n)
    ctx.gkr_gen_eq_evals(y, scale, out)
    assert np.array_equal(out.download(count=4 << n), oracle.gkr_gen_eq_evals(y, scale))


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("n_vars", [1, 6, 18])
def test_next_layer_matches_oracle(ctx, kind, n_vars):
    _say("next layer", kind, n_vars)
    rng = np.random.default_rng(kind * 100 + n_vars)
    n = 1 << n_vars
    den = rand_q(rng, n)
    num = rand_q(rng, n) if kind == 1 else rand_m(rng, n) if kind == 2 else None
    dd = dev(ctx, den)
    dn = dev(ctx, num) if num is not None else None
    on = ctx.alloc(16 * (n // 2)) if kind != 0 else None
    od = ctx.alloc(16 * (n // 2))
    ctx.gkr_next_layer(kind, dn, dd, n_vars, on, od)
    rn, rd = oracle.gkr_next_layer(kind, num, den, n_vars)
    assert np.array_equal(od.download(count=2 * n), rd)
    if kind != 0:
        assert np.array_equal(on.download(count=2 * n), rn)


@pytest.mark.parametrize("base", [False, True])
@pytest.mark.parametrize("n_vars", [1, 5, 18])
def test_mle_fix_first_variable_matches_oracle(ctx, base, n_vars):
    _say("fix first", base, n_vars)
    rng = np.random.default_rng(7 * n_vars + base)
    n = 1 << n_vars
    vals = rand_m(rng, n) if base else rand_q(rng, n)
    r = rand_q(rng, 1)
    out = ctx.alloc(16 * (n // 2))
    ctx.mle_fix_first_variable(base, dev(ctx, vals), n_vars, r, out)
    assert np.array_equal(out.download(count=2 * n), oracle.mle_fix_first(base, vals, n_vars, r))


def test_sum_as_poly_matches_definition(ctx):
    """First-round polynomial of a generic LogUp layer (6 variables, eq point y of 5 coordinates):
    r(t) = sum_x eq((t, x), y) * p(t, x) with p = fraction numerator + lambda * denominator of the pair
    (2i, 2i+1), values extended linearly in t (gkr_circuit.zig:390-394).  r(0) and r(2) are brute forced
    in Python; r(0) + r(1) = claim holds by construction."""
    import mathref as mr
    from gkr_ref import eq as eq_fn, horner, qm
    rng = np.random.default_rng(77)
    nv = 6
    n = 1 << nv
    num, den = rand_q(rng, n), rand_q(rng, n)
    y = rand_q(rng, nv - 1)
    lam = rand_q(rng, 1)
    claim = rand_q(rng, 1)
    ys = [tuple(int(v) for v in y[4 * j:4 * j + 4]) for j in range(nv - 1)]
    scale = mr.qsub((1, 0, 0, 0), ys[0])
    eq = oracle.gkr_gen_eq_evals(y[4:], np.array(scale, np.uint32))
    one = np.array([1, 0, 0, 0], np.uint32)
    coeffs = ctx.gkr_sum_as_poly_in_first_variable(1, dev(ctx, eq), dev(ctx, num), dev(ctx, den), nv, lam, claim,
                                                   y, one)
    cs = [tuple(int(v) for v in coeffs[4 * i:4 * i + 4]) for i in range(coeffs.size // 4)]
    assert 0 < len(cs) <= 4
    q = lambda arr, i: tuple(int(v) for v in arr[4 * i:4 * i + 4])  # noqa: E731
    L = tuple(int(v) for v in lam)
    nt = n // 4
    for t in (0, 2):
        acc = (0, 0, 0, 0)
        for x in range(nt):
            vals = []
            for off in (0, 1):
                v0n, v0d = q(num, 2 * x + off), q(den, 2 * x + off)
                v1n, v1d = q(num, 2 * (nt + x) + off), q(den, 2 * (nt + x) + off)
                if t == 0:
                    vals.append((v0n, v0d))
                else:
                    vals.append((mr.qsub(mr.qadd(v1n, v1n), v0n), mr.qsub(mr.qadd(v1d, v1d), v0d)))
            (n0, d0), (n1, d1) = vals
            p = mr.qadd(mr.qadd(mr.qmul(d1, n0), mr.qmul(d0, n1)), mr.qmul(L, mr.qmul(d0, d1)))
            bits = [qm(t)] + [qm((x >> (nv - 3 - j)) & 1) for j in range(nv - 2)]
            acc = mr.qadd(acc, mr.qmul(eq_fn(bits, ys), p))
        assert horner(cs, qm(t)) == acc
    assert mr.qadd(horner(cs, qm(0)), horner(cs, qm(1))) == tuple(int(v) for v in claim)


def _proof_case(ctx, shape):
    rng = np.random.default_rng({"gp": 1, "generic": 2, "mixed": 3, "xor12": 4, "xor18": 5, "gp_big": 6,
                                 "mixed_big": 7, "eight": 8, "one_var": 9, "edge8": 10, "edge9": 11}[shape])
    if shape == "gp":
        return [(0, None, rand_q(rng, 32), 5)]
    if shape == "gp_big":  # rounds above 2^8 values run over many workgroups
        return [(0, None, rand_q(rng, 1 << 11), 11)]
    if shape == "mixed_big":
        return [(1, rand_q(rng, 1 << 10), rand_q(rng, 1 << 10), 10), (0, None, rand_q(rng, 1 << 9), 9),
                (3, None, rand_q(rng, 1 << 8), 8), (2, rand_m(rng, 1 << 7, 1000), rand_q(rng, 1 << 7), 7)]
    if shape == "eight":  # the largest batch
        kinds = [0, 1, 2, 3, 1, 0, 3, 2]
        out = []
        for k, nv in zip(kinds, [9, 1, 4, 9, 2, 6, 3, 8]):
            num = rand_q(rng, 1 << nv) if k == 1 else (rand_m(rng, 1 << nv, 7) if k == 2 else None)
            out.append((k, num, rand_q(rng, 1 << nv), nv))
        return out
    if shape == "one_var":
        return [(1, rand_q(rng, 2), rand_q(rng, 2), 1)]
    if shape in ("edge8", "edge9"):  # around the one-workgroup threshold
        nv = 8 if shape == "edge8" else 9
        return [(1, rand_q(rng, 1 << nv), rand_q(rng, 1 << nv), nv), (3, None, rand_q(rng, 1 << nv), nv)]
    if shape == "generic":
        return [(1, rand_q(rng, 256), rand_q(rng, 256), 8)]
    if shape == "mixed":
        return [(3, None, rand_q(rng, 64), 6), (2, rand_m(rng, 16, 50), rand_q(rng, 16), 4),
                (0, None, rand_q(rng, 8), 3)]
    log = 12 if shape == "xor12" else 18
    a, b, c, m = oracle.xor_trace(log, 1234 + log)
    z, alpha = rand_q(rng, 1), rand_q(rng, 1)
    dl = oracle.logup_denominators(a, b, c, z, alpha)
    dt = oracle.logup_denominators(None, None, None, z, alpha)
    negm = ((P - m.astype(np.int64)) % P).astype(np.uint32)
    return [(3, None, dl, log), (2, negm, dt, 16)]


def _locate(words, n_layers, idx):
    """Names the part of a GKR proof word stream that holds word idx."""
    p = 0
    for layer in range(n_layers):
        nr = int(words[p])
        if idx == p:
            return "layer %d round count" % layer
        p += 1
        for rnd in range(nr):
            ln = int(words[p])
            if p <= idx < p + 1 + 4 * ln:
                return "layer %d round %d word %d" % (layer, rnd, idx - p)
            p += 1 + 4 * ln
        nm = int(words[p])
        if idx == p:
            return "layer %d mask count" % layer
        p += 1
        for _ in range(nm):
            nc = int(words[p + 1])
            if p <= idx < p + 2 + 8 * nc:
                return "layer %d mask of instance %d word %d" % (layer, int(words[p]), idx - p)
            p += 2 + 8 * nc
    return "tail (outputs, ood, claims, digest) word %d" % (idx - p)


@pytest.mark.parametrize("shape", ["gp", "generic", "mixed", "one_var", "edge8", "edge9", "gp_big", "mixed_big",
                                   "eight", "xor12", "xor18"])
def test_gkr_prove_batch_matches_oracle(ctx, shape):
    _say("gkr prove", shape)
    insts = _proof_case(ctx, shape)
    digest = bytes(range(32, 64))
    keep = []
    dinst = []
    for kind, num, den, nv in insts:
        dn = dev(ctx, num) if num is not None else None
        dd = dev(ctx, den)
        keep += [dn, dd]
        dinst.append((kind, dn, dd, nv))
    words = ctx.gkr_prove_batch(dinst, digest)
    ref = oracle.gkr_prove(insts, digest)
    assert words.size == ref.size
    diff = np.nonzero(words != ref)[0]
    where = _locate(ref, max(nv for _, _, _, nv in insts), int(diff[0])) if diff.size else ""
    assert diff.size == 0, "first difference at word %d: %s" % (int(diff[0]), where)


def test_xor_logup_balances_in_cumulative_columns_and_gkr(ctx):
    """The 2^18-row XOR lookup: lookup and table claimed sums cancel, and the GKR output fractions
    (num/den per instance) equal the cumulative columns' claimed sums."""
    import mathref as mr
    log = 18
    n = 1 << log
    rng = np.random.default_rng(11)
    z, alpha = rand_q(rng, 1), rand_q(rng, 1)
    bufs = [ctx.alloc(4 * n) for _ in range(3)]
    dm = ctx.alloc(4 << 16)
    ctx.xor_trace(log, 99, *bufs, dm)
    dl, dt = ctx.alloc(16 * n), ctx.alloc(16 << 16)
    ctx.logup_denominators(*bufs, n, z, alpha, dl)
    ctx.logup_denominators(None, None, None, 1 << 16, z, alpha, dt)
    m = dm.download(count=1 << 16)
    negm = dev(ctx, ((P - m.astype(np.int64)) % P).astype(np.uint32))
    sl, st = ctx.alloc(16 * n), ctx.alloc(16 << 16)
    cl = ctx.logup_cumulative(2, None, dl, n, sl)
    ct = ctx.logup_cumulative(1, negm, dt, 1 << 16, st)
    assert all((int(x) + int(y)) % P == 0 for x, y in zip(cl, ct))
    words = ctx.gkr_prove_batch([(3, None, dl, log), (2, negm, dt, 16)])
    ref = oracle.gkr_prove([(3, None, dl.download(count=4 * n), log),
                            (2, negm.download(count=1 << 16), dt.download(count=4 << 16), 16)])
    assert np.array_equal(words, ref)
    # output claims: [n, values] per instance after the layer section; num/den per instance
    tail = _output_claims(words, [log, 16])
    fr = [mr.qmul(nu, mr.qinv(de)) for nu, de in tail]
    assert fr[0] == tuple(int(v) for v in cl) and fr[1] == tuple(int(v) for v in ct)


def _output_claims(words, n_vars):
    """Skip the per-layer section of the proof stream and return (num, den) per instance."""
    pos = 0
    for _ in range(max(n_vars)):
        nr = int(words[pos])
        pos += 1
        for _ in range(nr):
            pos += 1 + 4 * int(words[pos])
        nm = int(words[pos])
        pos += 1
        for _ in range(nm):
            pos += 2 + 8 * int(words[pos + 1])
    out = []
    for _ in n_vars:
        k = int(words[pos])
        vals = [tuple(int(v) for v in words[pos + 1 + 4 * i:pos + 5 + 4 * i]) for i in range(k)]
        pos += 1 + 4 * k
        out.append((vals[0], vals[1]))
    return out


@pytest.mark.parametrize("log_n", [16, 18])
def test_logup_xor_pipeline_matches_oracle(ctx, log_n):
    """The whole LogUp workload: roots, claimed sums, sampled values, GKR proof and the final
    transcript digest equal the oracle's word for word."""
    _say("logup pipeline", log_n)
    words, _ = ctx.logup_xor_pipeline(log_n, 0x5EED + log_n)
    ref = oracle.logup_xor_pipeline(log_n, 0x5EED + log_n)
    assert words.size == ref.size
    assert np.array_equal(words, ref)
