"""Independent pure-Python restatement of the GKR batch *verifier* and of the
multilinear identities used to pin the oracle's LogUp/GKR functions.

Written from the verifier side of the reference (src/prover/lookups/
gkr_verifier.zig:139-300, sumcheck.zig:199-213, utils.zig:215-244) and the
Blake2s channel (src/core/channel/blake2s.zig:52-134), with hashlib as the
hash.  The prover's word stream is parsed, every sum-check round and layer
gate is checked, and the final claims are compared with direct multilinear
evaluations of the input layers.
"""
import hashlib
import struct

from mathref import P, qadd, qinv, qmul, qsub

ONE = (1, 0, 0, 0)
ZERO = (0, 0, 0, 0)


def qm(m):
    return (m % P, 0, 0, 0)


class Channel:
    """Blake2sChannel: digest, draw counter; mixFelts / drawSecureFelt."""

    def __init__(self, digest=bytes(32)):
        self.digest = digest
        self.n_draws = 0

    def mix_felts(self, felts):
        data = b"".join(struct.pack("<4I", *f) for f in felts)
        self.digest = hashlib.blake2s(self.digest + data).digest()
        self.n_draws = 0

    def draw_secure_felt(self):
        while True:
            h = hashlib.blake2s(self.digest + struct.pack("<I", self.n_draws) + b"\0").digest()
            self.n_draws += 1
            w = struct.unpack("<8I", h)
            if all(x < 2 * P for x in w):
                return tuple(x % P for x in w[:4])


def horner(coeffs, x):
    acc = ZERO
    for c in reversed(coeffs):
        acc = qadd(qmul(acc, x), c)
    return acc


def eq(xs, ys):
    out = ONE
    for x, y in zip(xs, ys):
        out = qmul(out, qadd(qmul(x, y), qmul(qsub(ONE, x), qsub(ONE, y))))
    return out


def fold(r, e0, e1):
    return qadd(qmul(r, qsub(e1, e0)), e0)


def mle_eval(values, point):
    """Multilinear extension at `point`; point[0] selects the upper half."""
    buf = list(values)
    for r in point:
        h = len(buf) // 2
        buf = [fold(r, buf[i], buf[i + h]) for i in range(h)]
    return buf[0]


def mle_eval_direct(values, point):
    """sum_x f(x) * eq(x, point) with x's most significant bit paired with point[0]."""
    n = len(point)
    acc = ZERO
    for i, v in enumerate(values):
        bits = [qm((i >> (n - 1 - j)) & 1) for j in range(n)]
        acc = qadd(acc, qmul(v, eq(bits, point)))
    return acc


def verify(words, instances, digest=bytes(32)):
    """instances: list of (gate 'logup'|'gp', n_vars).  Returns (ood, claims_to_verify, final digest)
    after checking every round; raises AssertionError on any failure."""
    n_inst = len(instances)
    n_layers = max(nv for _, nv in instances)
    pos = 0

    def u():
        nonlocal pos
        assert pos < len(words), "truncated proof stream"
        v = int(words[pos])
        pos += 1
        return v

    def cnt(limit=64):
        v = u()
        assert v <= limit, "malformed count in the proof stream"
        return v

    def q():
        nonlocal pos
        assert pos + 4 <= len(words), "truncated proof stream"
        v = tuple(int(x) for x in words[pos:pos + 4])
        pos += 4
        return v

    layers = []
    for _ in range(n_layers):
        rounds = []
        for _ in range(cnt()):
            ln = cnt(4)
            rounds.append([q() for _ in range(ln)])
        masks = {}
        for _ in range(cnt()):
            inst = cnt(n_inst - 1)
            ncols = cnt(2)
            flat = [q() for _ in range(2 * ncols)]
            masks[inst] = [(flat[2 * c], flat[2 * c + 1]) for c in range(ncols)]
        layers.append((rounds, masks))
    outputs = []
    for _ in range(n_inst):
        outputs.append([q() for _ in range(cnt(2))])
    ood_stream = [q() for _ in range(cnt())]
    final_claims = []
    for _ in range(n_inst):
        final_claims.append([q() for _ in range(cnt(2))])
        u()
    final_digest = b"".join(struct.pack("<I", u()) for _ in range(8))
    assert pos == len(words), "trailing words in the proof stream"

    ch = Channel(digest)
    ood = []
    claims = [None] * n_inst
    for layer, (rounds, masks) in enumerate(layers):
        remaining = n_layers - layer
        for k, (_, nv) in enumerate(instances):
            if nv == remaining:
                claims[k] = list(outputs[k])
        for k in range(n_inst):
            if claims[k] is not None:
                ch.mix_felts(claims[k])
        sc_alpha = ch.draw_secure_felt()
        lam = ch.draw_secure_felt()
        sc_claims, sc_inst = [], []
        for k, (_, nv) in enumerate(instances):
            if claims[k] is not None:
                unused = n_layers - nv
                sc_claims.append(qmul(horner(claims[k], lam), qm(pow(2, unused, P))))
                sc_inst.append(k)
        claim = horner(sc_claims, sc_alpha)
        assignment = []
        assert len(rounds) == layer, "sum-check has one round per variable of the layer"
        for poly in rounds:
            assert len(poly) <= 4, "round polynomial degree"
            assert qadd(horner(poly, ZERO), horner(poly, ONE)) == claim, "sum-check round sum"
            ch.mix_felts(poly)
            r = ch.draw_secure_felt()
            claim = horner(poly, r)
            assignment.append(r)
        evals = []
        for k in sc_inst:
            gate, nv = instances[k]
            unused = n_layers - nv
            mask = masks[k]
            if gate == "logup":
                (n0, n1), (d0, d1) = mask
                out = [qadd(qmul(d1, n0), qmul(d0, n1)), qmul(d0, d1)]
            else:
                (v0, v1), = mask
                out = [qmul(v0, v1)]
            evals.append(qmul(eq(ood[unused:], assignment[unused:]), horner(out, lam)))
        assert horner(evals, sc_alpha) == claim, "layer gate check"
        for k in sc_inst:
            flat = [v for col in masks[k] for v in col]
            ch.mix_felts(flat)
        challenge = ch.draw_secure_felt()
        ood = assignment + [challenge]
        for k in sc_inst:
            claims[k] = [fold(challenge, c0, c1) for c0, c1 in masks[k]]
    assert ood == ood_stream, "ood point"
    assert claims == final_claims, "claims to verify"
    assert ch.digest == final_digest, "final channel digest"
    return ood, claims, ch.digest
