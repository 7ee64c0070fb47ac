"""Pins the C oracle (oracle/stwo_oracle.c) before it is trusted as the checker.

1. End to end: the oracle reproduces, byte for byte, every committed golden
   wide-Fibonacci proof of the reference (5x16 pow0 interop proof, 10x8, 14x32,
   16x64, 18x100, 20x100 functional), both Poseidon AIR proofs (log_n_instances
   10 and 13, 1264 columns) and both periodic XOR AIR proofs (log_size 14 and 16).  Each proof pins the trace LDE + Merkle
   root, the composition root, sampled values, quotients, all FRI layer roots
   and witnesses, the last layer, the PoW nonce and the decommitments.
2. Primitives, against closed forms / independent restatements:
   fields vs Python big ints, Blake2s vs hashlib and the reference KATs,
   circle group and domains vs mathref, FFT vs direct O(n^2) evaluation of the
   circle basis, twiddle inverses, fold and PoW properties.
"""
import hashlib
import random

import numpy as np
import pytest

import mathref as mr
import oracle
from conftest import (golden_kwargs, load_goldens, load_poseidon_goldens, load_xor_goldens, poseidon_kwargs,
                      xor_kwargs)

GOLDENS = load_goldens()
POSEIDON_GOLDENS = load_poseidon_goldens()
XOR_GOLDENS = load_xor_goldens()


@pytest.mark.parametrize("g", GOLDENS, ids=[g["name"] for g in GOLDENS])
def test_oracle_reproduces_golden_proof(g):
    proof = oracle.prove_wide_fibonacci(**golden_kwargs(g))
    assert hashlib.sha256(proof).hexdigest() == g["sha256"]
    assert proof == g["proof"]


@pytest.mark.parametrize("g", POSEIDON_GOLDENS, ids=[g["name"] for g in POSEIDON_GOLDENS])
def test_oracle_reproduces_poseidon_golden_proof(g):
    # src/examples/poseidon.zig: 1264-column trace, constant composition, statement log_n_instances
    proof = oracle.prove_poseidon(**poseidon_kwargs(g))
    assert hashlib.sha256(proof).hexdigest() == g["sha256"]
    assert proof == g["proof"]


@pytest.mark.parametrize("g", XOR_GOLDENS, ids=[g["name"] for g in XOR_GOLDENS])
def test_oracle_reproduces_xor_golden_proof(g):
    # src/examples/xor.zig: two preprocessed indicator columns (sampled like every column),
    # one main column, statement mixU32s(log_size, log_step) + mixU64(offset)
    proof = oracle.prove_xor(**xor_kwargs(g))
    assert hashlib.sha256(proof).hexdigest() == g["sha256"]
    assert proof == g["proof"]


def test_xor_example_columns_are_indicators():
    # IsFirst: only storage index 0; IsStepWithOffset: one row in 2^log_step; main: every
    # even coset row, each at its bit-reversed circle-domain position (a permutation).
    log, step, off = 6, 2, 3
    first, stepc, main = oracle.xor_example_columns(log, step, off)
    n = 1 << log
    assert first[0] == 1 and first.sum() == 1
    assert stepc.sum() == n >> step and main.sum() == n // 2

    def pos(i):  # coset index -> bit-reversed circle-domain index, from the definitions
        ci = i // 2 if i % 2 == 0 else (2 * n - i) // 2
        return int(format(ci, "0%db" % log)[::-1], 2)

    assert sorted(pos(i) for i in range(n)) == list(range(n))
    assert {pos(i) for i in range(off % (1 << step), n, 1 << step)} == set(np.nonzero(stepc)[0].tolist())
    assert {pos(i) for i in range(0, n, 2)} == set(np.nonzero(main)[0].tolist())


def test_poseidon_trace_inputs_and_first_round():
    # input.zig:121-133: instance r of row k starts from the state (16k + i + r)_i; the first
    # external round of the all-zero-offset state is checked against big-int arithmetic.
    t = oracle.poseidon_trace(2)
    assert t.shape == (1264, 4)
    for rep in range(8):
        for k in range(4):
            assert [int(t[158 * rep + i, k]) for i in range(16)] == [16 * k + i + rep for i in range(16)]
    P = (1 << 31) - 1
    s = [(i + 1234 + i) % P for i in range(16)]  # row 0, instance 0: state i, round constant 1234 + i

    def m4(x):
        t0, t1 = x[0] + x[1], x[2] + x[3]
        t2, t3 = 2 * x[1] + t1, 2 * x[3] + t0
        t4, t5 = 4 * t1 + t3, 4 * t0 + t2
        return [t3 + t5, t5, t2 + t4, t4]

    s = sum((m4(s[4 * i:4 * i + 4]) for i in range(4)), [])
    sums = [s[j] + s[j + 4] + s[j + 8] + s[j + 12] for j in range(4)]
    s = [pow((s[i] + sums[i % 4]) % P, 5, P) for i in range(16)]
    assert [int(v) for v in t[16:32, 0]] == s


def test_golden_20x100_is_the_receipt_proof():
    g = [g for g in GOLDENS if g["name"] == "wide_fibonacci_log20_seq100_pow10"][0]
    # receipt.json canonical_proof_sha256 of conformance/evidence/metrics-v2/w1-log20x100
    assert g["sha256"] == "e6609d0564a47192212bec7973e2660c2eea88bef90c573c3df09569cc3c7e86"
    assert g["bytes"] == 86383


def test_m31_mul_inv_bruteforce():
    rng = random.Random(1)
    for _ in range(2000):
        a, b = rng.randrange(mr.P), rng.randrange(mr.P)
        assert oracle.m31_mul(a, b) == a * b % mr.P
    for a in [1, 2, mr.P - 1, 12345] + [rng.randrange(1, mr.P) for _ in range(200)]:
        assert oracle.m31_inv(a) * a % mr.P == 1


def test_qm31_mul_inv_vs_definition():
    rng = random.Random(2)
    for _ in range(300):
        x = tuple(rng.randrange(mr.P) for _ in range(4))
        y = tuple(rng.randrange(mr.P) for _ in range(4))
        assert tuple(oracle.qm31_mul(x, y)) == mr.qmul(x, y)
    for _ in range(20):
        x = tuple(rng.randrange(mr.P) for _ in range(4))
        assert tuple(oracle.qm31_inv(x)) == mr.qinv(x)
        assert mr.qmul(x, tuple(oracle.qm31_inv(x))) == (1, 0, 0, 0)


def test_blake2s_vs_hashlib():
    rng = random.Random(3)
    for n in [0, 1, 3, 31, 32, 37, 40, 52, 63, 64, 65, 80, 96, 127, 128, 129, 464, 1000]:
        data = bytes(rng.randrange(256) for _ in range(n))
        assert oracle.blake2s(data) == hashlib.blake2s(data).digest()
    # src/core/vcs/blake2_hash.zig:297-302
    assert oracle.blake2s(b"a").hex() == "4a0d129873403037c2cd9b9048203687f6233fb6738956e0349bd4320fec3e90"


def test_merkle_hash_kats():
    # src/core/vcs_lifted/blake2_merkle.zig:364-384: prefixed empty leaf and parent
    root, layers = oracle.merkle_commit([])
    assert root.hex() == "2a133e150238721921d1ea772882979c810f85f2849099b9d3415a8619d85fad"
    node = hashlib.blake2s(b"node" + bytes(60) + b"\x01" * 32 + b"\x02" * 32).hexdigest()
    assert node == "24c36247c66cc7b145aee79ccff9d5e3e596a8e13e86d13942f61464364fa53c"
    # plain hashes (blake2_merkle.zig:339-361) pin the byte layout of leaves
    assert hashlib.blake2s(np.array([1, 2, 3], dtype="<u4").tobytes()).hexdigest() == \
        "9549ce18083a73c794b0fe338635a4ab0950333ecb2e3f1945adf7db5f0ef4d0"
    assert hashlib.blake2s(b"\x01" * 32 + b"\x02" * 32).hexdigest() == \
        "280569932378c99f448df37e893f062fab951bea53515634b7875ae51e1954e7"


def _merkle_ref(columns):
    """Lifted Merkle from the definition (leaves.zig:24-111, layers.zig:99-140)."""
    order = sorted(range(len(columns)), key=lambda i: (len(columns[i]), i))
    L = max(len(c) for c in columns).bit_length() - 1
    leaves = []
    for p in range(1 << L):
        msg = b"leaf" + bytes(60)
        for i in order:
            lg = len(columns[i]).bit_length() - 1
            idx = ((p >> (L - lg + 1)) << 1) + (p & 1)
            msg += int(columns[i][idx]).to_bytes(4, "little")
        leaves.append(hashlib.blake2s(msg).digest())
    layers = [leaves]
    while len(layers[-1]) > 1:
        prev = layers[-1]
        layers.append([hashlib.blake2s(b"node" + bytes(60) + prev[2 * i] + prev[2 * i + 1]).digest()
                       for i in range(len(prev) // 2)])
    return layers[-1][0], layers


def test_merkle_lifted_mixed_sizes_vs_definition():
    rng = np.random.default_rng(4)
    cols = [rng.integers(0, mr.P, 1 << k, dtype=np.uint32) for k in (3, 1, 5, 3, 2, 5)]
    root, layers = oracle.merkle_commit(cols)
    ref_root, ref_layers = _merkle_ref(cols)
    assert root == ref_root
    assert layers[5] == b"".join(ref_layers[0])
    # 17 columns: leaf message spans two compressions plus a partial one
    cols = [rng.integers(0, mr.P, 16, dtype=np.uint32) for _ in range(17)]
    assert oracle.merkle_commit(cols)[0] == _merkle_ref(cols)[0]
    # all-constant columns take a special path in the reference; same result
    cols = [np.full(8, 7, dtype=np.uint32), np.full(4, 9, dtype=np.uint32)]
    assert oracle.merkle_commit(cols)[0] == _merkle_ref(cols)[0]


def test_circle_points_vs_definition():
    rng = random.Random(5)
    assert oracle.circle_point_at_index(0) == (1, 0)
    assert oracle.circle_point_at_index(1 << 30) == (mr.P - 1, 0)  # element of order 2
    for _ in range(50):
        i = rng.randrange(1 << 31)
        x, y = oracle.circle_point_at_index(i)
        assert (x * x + y * y) % mr.P == 1
        assert (x, y) == mr.point_at_index(i)
    for log in (1, 2, 3, 7, 12):
        for i in range(0, 1 << log, max(1, (1 << log) // 16)):
            assert oracle.circle_domain_point(log, i) == mr.point_at_index(mr.domain_index(log, i))


def test_twiddles_inverse_and_tower_suffix():
    for log in (3, 5, 9):
        tw, itw = oracle.twiddles(log)
        assert all(int(a) * int(b) % mr.P == 1 for a, b in zip(tw, itw))
        assert tw[-1] == 1
        # first layer = x of the half coset's first half, bit-reversed (twiddles.zig:67-91)
        half = log - 1
        init, step = 1 << (31 - half - 2), 1 << (31 - half)
        n0 = 1 << (half - 1)
        for j in range(n0):
            nat = mr.bitrev(j, half - 1)
            assert tw[j] == mr.point_at_index(init + step * nat)[0]
        # tower: the log-(k) tree is the suffix of the log-(k+2) tree (twiddle_tower.zig:66-87)
        big, _ = oracle.twiddles(log + 2)
        assert list(big[len(big) - len(tw):]) == list(tw)


@pytest.mark.parametrize("log", [1, 2, 3, 4, 5, 6])
def test_fft_matches_direct_circle_basis_evaluation(log):
    rng = np.random.default_rng(log)
    coeffs = rng.integers(0, mr.P, 1 << log, dtype=np.uint32)
    evals = oracle.circle_evaluate(coeffs, log)
    for r in range(1 << log):
        x, y = mr.point_at_index(mr.domain_index(log, mr.bitrev(r, log)))
        assert int(evals[r]) == mr.circle_basis_eval([int(c) for c in coeffs], x, y)
    assert np.array_equal(oracle.circle_interpolate(evals, log), coeffs)


def test_fft_roundtrip_and_lde():
    rng = np.random.default_rng(7)
    for log in (7, 10, 13):
        v = rng.integers(0, mr.P, 1 << log, dtype=np.uint32)
        c = oracle.circle_interpolate(v, log)
        assert np.array_equal(oracle.circle_evaluate(c, log), v)
        # LDE restricted to the original domain: CanonicCoset(log+1) contains
        # CanonicCoset(log)'s points? No: check low degree instead -- interpolating
        # the LDE gives the coefficients followed by zeros.
        lde = oracle.circle_lde(c, log, 2)
        c2 = oracle.circle_interpolate(lde, log + 2)
        assert np.array_equal(c2[: 1 << log], c) and not c2[1 << log:].any()


def test_eval_at_point_vs_definition():
    rng = np.random.default_rng(8)
    coeffs = rng.integers(0, mr.P, 32, dtype=np.uint32)
    x, y = mr.point_at_index(12345)
    pt8 = [x, 0, 0, 0, y, 0, 0, 0]
    assert oracle.eval_at_point(coeffs, 5, pt8) == [mr.circle_basis_eval([int(c) for c in coeffs], x, y), 0, 0, 0]


def test_fold_properties():
    rng = np.random.default_rng(9)
    alpha = [int(v) for v in rng.integers(0, mr.P, 4)]
    # folding a constant line evaluation c gives 2c for any alpha
    c = [5, 6, 7, 8]
    src = np.tile(np.array(c, dtype=np.uint32), 16)
    out = oracle.fold_line(src, 4, 1 << (31 - 6), alpha).reshape(-1, 4)
    assert (out == np.array([2 * v for v in c], dtype=np.uint32)).all()
    # linearity in the source
    a = rng.integers(0, mr.P, 64, dtype=np.uint32)
    b = rng.integers(0, mr.P, 64, dtype=np.uint32)
    s = ((a.astype(np.uint64) + b) % mr.P).astype(np.uint32)
    fa, fb, fs = (oracle.fold_line(v, 4, 1 << (31 - 6), alpha) for v in (a, b, s))
    assert np.array_equal(((fa.astype(np.uint64) + fb) % mr.P).astype(np.uint32), fs)


def test_pow_grind_is_lowest_valid_nonce():
    digest = hashlib.blake2s(b"pow").digest()
    bits = 8
    nonce = oracle.pow_grind(digest, bits)
    prefix = hashlib.blake2s((0x12345678).to_bytes(4, "little") + bytes(12) + digest + bits.to_bytes(4, "little")).digest()

    def tz(n):
        h = int.from_bytes(hashlib.blake2s(prefix + n.to_bytes(8, "little")).digest()[:16], "little")
        return 128 if h == 0 else (h & -h).bit_length() - 1

    assert tz(nonce) >= bits
    assert all(tz(n) < bits for n in range(nonce))


def test_widefib_trace_recurrence_and_seeds():
    t = oracle.widefib_trace(5, 6, mode=0)
    n = 32
    for s in range(n):
        ci = mr.bitrev(s, 5)
        row = ci * 2 if ci < n // 2 else (n - 1 - ci) * 2 + 1
        assert t[0][s] == 1 and t[1][s] == row
    for c in range(2, 6):
        assert np.array_equal(t[c], ((t[c - 2].astype(np.uint64) ** 2 + t[c - 1].astype(np.uint64) ** 2) % mr.P))
    u = oracle.widefib_trace(5, 3, mode=1, seed=42)
    assert len(set(u[0].tolist())) > 20 and (u < mr.P).all()
