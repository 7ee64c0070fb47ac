"""Pins the oracle's LogUp / GKR restatement (CPU only).

Each check is independent of the oracle's formulas: direct fraction sums with
Python big ints, the multilinear identity sum_x f(x) eq(x, r), a restatement of
the GKR verifier that replays the Blake2s transcript with hashlib, and the
balance of the XOR lookup (lookups minus table multiplicities sum to zero).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import oracle  # noqa: E402
from gkr_ref import ONE, ZERO, eq, mle_eval, mle_eval_direct, qm, verify  # noqa: E402
from mathref import P, qadd, qinv, qmul  # noqa: E402


def q_of(words, i):
    return tuple(int(x) for x in words[4 * i:4 * i + 4])


def rand_q(rng, n):
    return rng.integers(0, P, size=4 * n, dtype=np.uint64).astype(np.uint32)


def frac_sum(nums, dens):
    acc = ZERO
    for n, d in zip(nums, dens):
        acc = qadd(acc, qmul(n, qinv(d)))
    return acc


def test_eq_evals_match_direct_eq():
    rng = np.random.default_rng(1)
    y = rand_q(rng, 3)
    scale = rand_q(rng, 1)
    out = oracle.gkr_gen_eq_evals(y, scale)
    ys = [q_of(y, j) for j in range(3)]
    for i in range(8):
        bits = [qm((i >> (2 - j)) & 1) for j in range(3)]  # y[0] selects the upper half
        assert q_of(out, i) == qmul(q_of(scale, 0), eq(bits, ys))


@pytest.mark.parametrize("kind", [oracle.GKR_GRAND_PRODUCT, oracle.GKR_LOGUP_GENERIC,
                                  oracle.GKR_LOGUP_MULTIPLICITIES, oracle.GKR_LOGUP_SINGLES])
def test_next_layer_to_output_equals_direct_sum(kind):
    rng = np.random.default_rng(10 + kind)
    nv = 6
    n = 1 << nv
    den = rand_q(rng, n)
    num = None
    if kind == oracle.GKR_LOGUP_GENERIC:
        num = rand_q(rng, n)
    elif kind == oracle.GKR_LOGUP_MULTIPLICITIES:
        num = rng.integers(0, P, size=n, dtype=np.uint64).astype(np.uint32)
    if kind == oracle.GKR_GRAND_PRODUCT:
        expect = ONE
        for i in range(n):
            expect = qmul(expect, q_of(den, i))
    else:
        nums = [q_of(num, i) if kind == oracle.GKR_LOGUP_GENERIC else
                qm(int(num[i])) if kind == oracle.GKR_LOGUP_MULTIPLICITIES else ONE for i in range(n)]
        expect = frac_sum(nums, [q_of(den, i) for i in range(n)])
    k = kind
    for v in range(nv, 0, -1):
        num, den = oracle.gkr_next_layer(k, num, den, v)
        if k != oracle.GKR_GRAND_PRODUCT:
            k = oracle.GKR_LOGUP_GENERIC
    if kind == oracle.GKR_GRAND_PRODUCT:
        assert q_of(den, 0) == expect
    else:
        assert qmul(q_of(num, 0), qinv(q_of(den, 0))) == expect


@pytest.mark.parametrize("base", [False, True])
def test_fix_first_variable_matches_multilinear_identity(base):
    rng = np.random.default_rng(20 + base)
    nv = 5
    n = 1 << nv
    vals = rng.integers(0, P, size=n if base else 4 * n, dtype=np.uint64).astype(np.uint32)
    point = [q_of(rand_q(rng, 1), 0) for _ in range(nv)]
    cur, b = vals, base
    for j, r in enumerate(point):
        cur = oracle.mle_fix_first(b, cur, nv - j, np.array(r, np.uint32))
        b = False
    f = [qm(int(v)) for v in vals] if base else [q_of(vals, i) for i in range(n)]
    assert q_of(cur, 0) == mle_eval_direct(f, point)


def test_cumulative_column_matches_prefix_fractions():
    rng = np.random.default_rng(3)
    n = 64
    num, den = rand_q(rng, n), rand_q(rng, n)
    sums, claimed = oracle.logup_cumulative(0, num, den)
    acc = ZERO
    for i in range(n):
        acc = qadd(acc, qmul(q_of(num, i), qinv(q_of(den, i))))
        assert q_of(sums, i) == acc
    assert tuple(int(x) for x in claimed) == acc
    with pytest.raises(ZeroDivisionError):
        oracle.logup_cumulative(2, None, np.zeros(16, np.uint32))


def test_xor_trace_and_balanced_lookup_sum():
    a, b, c, m = oracle.xor_trace(12, 99)
    assert np.array_equal(c, a ^ b) and a.max() < 256 and b.max() < 256
    assert np.array_equal(m, np.bincount((a.astype(np.int64) << 8) | b, minlength=1 << 16).astype(np.uint32))
    z = np.array([123456, 7, 89, 1011], np.uint32)
    alpha = np.array([5, 55, 555, 5555], np.uint32)
    dl = oracle.logup_denominators(a, b, c, z, alpha)
    dt = oracle.logup_denominators(None, None, None, z, alpha)
    # a few denominators against the definition z - (a + b alpha + c alpha^2)
    al, al2 = tuple(int(x) for x in alpha), None
    al2 = qmul(al, al)
    for i in (0, 1, 4095):
        comb = qadd(qadd(qm(int(a[i])), qmul(al, qm(int(b[i])))), qmul(al2, qm(int(c[i]))))
        assert q_of(dl, i) == tuple((int(zz) - cc) % P for zz, cc in zip(z, comb))
    _, cl = oracle.logup_cumulative(2, None, dl)
    negm = ((P - m.astype(np.int64)) % P).astype(np.uint32)
    _, ct = oracle.logup_cumulative(1, negm, dt)
    assert all((int(x) + int(y)) % P == 0 for x, y in zip(cl, ct))


def _layer_values(kind, num, den, n):
    if kind == oracle.GKR_GRAND_PRODUCT:
        return [[q_of(den, i) for i in range(n)]]
    nums = [q_of(num, i) if kind == oracle.GKR_LOGUP_GENERIC else
            qm(int(num[i])) if kind == oracle.GKR_LOGUP_MULTIPLICITIES else ONE for i in range(n)]
    return [nums, [q_of(den, i) for i in range(n)]]


@pytest.mark.parametrize("shape", ["gp", "generic", "xor_small", "eight", "one_var"])
def test_gkr_proof_verifies_and_final_claims_are_input_mle_evals(shape):
    rng = np.random.default_rng({"gp": 5, "generic": 6, "xor_small": 7, "eight": 12, "one_var": 13}[shape])
    if shape == "gp":
        insts = [(oracle.GKR_GRAND_PRODUCT, None, rand_q(rng, 32), 5)]
    elif shape == "eight":  # every kind, eight instances of different sizes
        insts = []
        for k, nv in zip([0, 1, 2, 3, 1, 0, 3, 2], [5, 1, 4, 5, 2, 3, 3, 4]):
            num = rand_q(rng, 1 << nv) if k == 1 else (
                rng.integers(0, 7, size=1 << nv, dtype=np.uint64).astype(np.uint32) if k == 2 else None)
            insts.append((k, num, rand_q(rng, 1 << nv), nv))
    elif shape == "one_var":
        insts = [(oracle.GKR_LOGUP_GENERIC, rand_q(rng, 2), rand_q(rng, 2), 1)]
    elif shape == "generic":
        insts = [(oracle.GKR_LOGUP_GENERIC, rand_q(rng, 16), rand_q(rng, 16), 4)]
    else:
        mult = rng.integers(0, 9, size=16, dtype=np.uint64).astype(np.uint32)
        insts = [(oracle.GKR_LOGUP_SINGLES, None, rand_q(rng, 64), 6),
                 (oracle.GKR_LOGUP_MULTIPLICITIES, ((P - mult.astype(np.int64)) % P).astype(np.uint32),
                  rand_q(rng, 16), 4)]
    digest = bytes(range(32))
    words = oracle.gkr_prove(insts, digest)
    gates = [("gp" if k == oracle.GKR_GRAND_PRODUCT else "logup", nv) for k, _, _, nv in insts]
    ood, claims, _ = verify(words, gates, digest)
    n_layers = max(nv for _, _, _, nv in insts)
    for (kind, num, den, nv), cl in zip(insts, claims):
        point = ood[n_layers - nv:]
        for col, c in zip(_layer_values(kind, num, den, 1 << nv), cl):
            assert c == mle_eval(col, point)


def test_gkr_verifier_rejects_tampered_proof():
    rng = np.random.default_rng(8)
    insts = [(oracle.GKR_LOGUP_GENERIC, rand_q(rng, 16), rand_q(rng, 16), 4)]
    words = oracle.gkr_prove(insts)
    bad = words.copy()
    # stream: [0 rounds][1 mask][instance 0][2 columns][mask values...]: change the first mask value
    bad[4] = (int(bad[4]) + 1) % P
    with pytest.raises(AssertionError):
        verify(bad, [("logup", 4)])


def _np_qmul(x, y):
    """Vectorised QM31 product of (4, n) int64 arrays (u^2 = 2 + i)."""
    a0, a1, b0, b1 = x
    c0, c1, d0, d1 = y

    def cm(p0, p1, q0, q1):
        return (p0 * q0 - p1 * q1) % P, (p0 * q1 + p1 * q0) % P

    ac = cm(a0, a1, c0, c1)
    bd = cm(b0, b1, d0, d1)
    rbd = ((2 * bd[0] - bd[1]) % P, (bd[0] + 2 * bd[1]) % P)
    ad = cm(a0, a1, d0, d1)
    bc = cm(b0, b1, c0, c1)
    return np.array([(ac[0] + rbd[0]) % P, (ac[1] + rbd[1]) % P, (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P])


def _np_mle_eval(words, point):
    """MLE of a QM31 AoS column (uint32 words) at `point` (point[0] folds the upper half)."""
    v = words.reshape(-1, 4).T.astype(np.int64)
    for r in point:
        h = v.shape[1] // 2
        lo, hi = v[:, :h], v[:, h:]
        rr = np.array(r, np.int64).reshape(4, 1) * np.ones((1, h), np.int64)
        v = (_np_qmul(rr, (hi - lo) % P) + lo) % P
    return tuple(int(x) for x in v[:, 0])


def test_logup_pipeline_transcript_and_gkr_verify():
    """Replays the LogUp workload's transcript up to the GKR proof with hashlib, verifies the GKR
    proof embedded in the oracle's output, and checks its final claims against the lookup and
    table layers (denominators recomputed from the replayed z, alpha; numerators 1 and -m)."""
    import struct
    from gkr_ref import Channel, verify as gkr_verify
    log_n, seed = 16, 0x5EED + 16
    words = oracle.logup_xor_pipeline(log_n, seed)
    roots = [b"".join(struct.pack("<I", int(w)) for w in words[8 * r:8 * r + 8]) for r in range(3)]
    cl, ct = q_of(words[24:32], 0), q_of(words[24:32], 1)
    assert all((x + y) % P == 0 for x, y in zip(cl, ct))
    gn = int(words[24 + 8 + 48])
    gw = words[24 + 8 + 48 + 1:24 + 8 + 48 + 1 + gn]
    ch = Channel()
    ch.digest = __import__("hashlib").blake2s(ch.digest + struct.pack("<3I", log_n, seed & 0xffffffff,
                                                                       seed >> 32)).digest()
    ch.digest = __import__("hashlib").blake2s(ch.digest + roots[0]).digest()
    ch.n_draws = 0
    z, alpha = ch.draw_secure_felt(), ch.draw_secure_felt()
    ch.mix_felts([cl, ct])
    ch.digest = __import__("hashlib").blake2s(ch.digest + roots[1]).digest()
    ch.n_draws = 0
    ood, claims, _ = gkr_verify(gw, [("logup", log_n), ("logup", log_n)], ch.digest)
    a, b, c, m = oracle.xor_trace(log_n, seed)
    dl = oracle.logup_denominators(a, b, c, np.array(z, np.uint32), np.array(alpha, np.uint32))
    dt = oracle.logup_denominators(None, None, None, np.array(z, np.uint32), np.array(alpha, np.uint32),
                                   n=1 << log_n)
    assert claims[0][0] == ONE and claims[0][1] == _np_mle_eval(dl, ood)
    negm = np.zeros(4 << log_n, np.uint32)
    negm[0::4] = (P - m.astype(np.int64)) % P
    assert claims[1][0] == _np_mle_eval(negm, ood) and claims[1][1] == _np_mle_eval(dt, ood)
