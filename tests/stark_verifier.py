"""Independent pure-Python restatement of the reference STARK *verifier* for the proofs
this backend produces (wide-Fibonacci, Poseidon and periodic XOR example AIRs, and this
backend's XOR-lookup AIR with its LogUp interaction tree), used as
an acceptance oracle: a proof that no committed golden covers is checked here, not
only against the C oracle (test infrastructure; never on the product path).

Written from the verifier side of the reference with hashlib as Blake2s and Python
integers as M31/CM31/QM31:

  statement / transcript   src/examples/wide_fibonacci.zig:360-426, poseidon.zig:330-394,
                           xor.zig:352-419, 599-605; core/pcs/mod.zig:32-52 (mixInto)
  AIR at the OODS point    core/verifier.zig:24-107; core/proof.zig:37-69 (composition
                           split); core/air/accumulation.zig:9-28; core/constraints.zig:16-27
                           (coset vanishing); wide_fibonacci/component.zig:105-134
  PCS                      core/pcs/verifier.zig:59-178 (commit, verifyValues),
                           core/pcs/utils.zig:292-314 (preprocessed query positions)
  quotients at the queries core/pcs/quotients/fri_answers.zig:107-173, samples.zig:245-323,
                           row_evaluation.zig:36-43,120-141,506-521, constraints.zig:87-97
  FRI                      core/fri.zig:52-214 (commit, queries, decommit, last layer),
                           core/fri/folding.zig:20-235 (subset rebuild and folds),
                           core/queries.zig (draw, dedup, fold), core/poly/utils.zig:14-35
  lifted Merkle            core/vcs_lifted/verifier.zig:93-244, blake2_merkle.zig (prefixes)
  channel                  core/channel/blake2s.zig:52-166 (mix, draws, PoW)

verify_proof(...) returns None or raises VerificationError with the failing check.
Only fold_step 1 proofs are verified (the only kind this backend emits); any other
fold_step is rejected.
"""
import hashlib
import json
import struct

P = (1 << 31) - 1
ORDER = 1 << 31
GEN = (2, 1268011823)  # M31 circle generator (core/circle.zig:137-143)
LEAF_SEED = b"leaf" + bytes(60)
NODE_SEED = b"node" + bytes(60)
POW_PREFIX = 0x12345678
POSEIDON_COLUMNS, POSEIDON_COLUMNS_PER_REP = 1264, 158


class VerificationError(Exception):
    pass


def _check(cond, what):
    if not cond:
        raise VerificationError(what)


# ---------------------------------------------------------------- fields
# QM31 = (a, b, c, d) meaning (a + b i) + (c + d i) u, i^2 = -1, u^2 = 2 + i.
ZERO = (0, 0, 0, 0)
ONE = (1, 0, 0, 0)


def _cm(x, y):
    return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def qadd(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P, (x[2] + y[2]) % P, (x[3] + y[3]) % P)


def qsub(x, y):
    return ((x[0] - y[0]) % P, (x[1] - y[1]) % P, (x[2] - y[2]) % P, (x[3] - y[3]) % P)


def qneg(x):
    return tuple((-v) % P for v in x)


def qmul(x, y):
    a, b, c, d = (x[0], x[1]), (x[2], x[3]), (y[0], y[1]), (y[2], y[3])
    bd = _cm(b, d)
    rbd = _cm((2, 1), bd)
    ac = _cm(a, c)
    re = ((ac[0] + rbd[0]) % P, (ac[1] + rbd[1]) % P)
    ad, bc = _cm(a, d), _cm(b, c)
    return (re[0], re[1], (ad[0] + bc[0]) % P, (ad[1] + bc[1]) % P)


def qm(v):
    return (v % P, 0, 0, 0)


def qinv(x):
    _check(x != ZERO, "inverse of zero")
    # x^(p^4 - 2) by square and multiply
    r, b, e = ONE, x, P ** 4 - 2
    while e:
        if e & 1:
            r = qmul(r, b)
        b = qmul(b, b)
        e >>= 1
    return r


def conj(x):
    """complexConjugate: negate the u-part (qm31.zig:234-239)."""
    return (x[0], x[1], (-x[2]) % P, (-x[3]) % P)


def cm_inv(a):
    n = (a[0] * a[0] + a[1] * a[1]) % P
    _check(n != 0, "CM31 inverse of zero")
    ni = pow(n, P - 2, P)
    return (a[0] * ni % P, (-a[1]) * ni % P)


# ---------------------------------------------------------------- circle
def padd(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def point_at(idx):
    """generator^idx on the M31 circle (index arithmetic mod 2^31)."""
    r, b, k = (1, 0), GEN, idx % ORDER
    while k:
        if k & 1:
            r = padd(r, b)
        b = padd(b, b)
        k >>= 1
    return r


def subgroup_gen(log):
    return 1 << (31 - log)


def circle_domain_index(log, i):
    """CanonicCoset(log).circleDomain().indexAt(i): half coset halfOdds(log-1), then its negation."""
    half = 1 << (log - 1)
    init, step = subgroup_gen(log + 1), subgroup_gen(log - 1)
    if i < half:
        return (init + step * i) % ORDER
    return (-(init + step * (i - half))) % ORDER


def bitrev(i, n):
    r = 0
    for _ in range(n):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def qp_double_x(x):
    return qsub(qadd(qmul(x, x), qmul(x, x)), ONE)


def qp_add(p, q):
    return (qsub(qmul(p[0], q[0]), qmul(p[1], q[1])), qadd(qmul(p[0], q[1]), qmul(p[1], q[0])))


def coset_vanishing_qm(log_size, p):
    """cosetVanishing(CanonicCoset(log).coset(), p) (constraints.zig:16-27) for a QM31 point:
    x of p - initial + half_step, doubled log - 1 times.  CanonicCoset(log).coset() =
    Coset.odds(log): initial index 2^(31-log-1), step 2^(31-log), half step 2^(31-log-1)."""
    init = point_at(subgroup_gen(log_size + 1))
    half = point_at(subgroup_gen(log_size + 1))
    neg_init = (init[0], (-init[1]) % P)
    s = qp_add(p, (qm(neg_init[0]), qm(neg_init[1])))
    s = qp_add(s, (qm(half[0]), qm(half[1])))
    x = s[0]
    for _ in range(1, log_size):
        x = qp_double_x(x)
    return x


# ---------------------------------------------------------------- channel
class Channel:
    """Blake2sChannel (core/channel/blake2s.zig:52-166)."""

    def __init__(self):
        self.digest = bytes(32)
        self.n_draws = 0

    def mix_bytes(self, data):
        self.digest = hashlib.blake2s(self.digest + data).digest()
        self.n_draws = 0

    def mix_felts(self, felts):
        self.mix_bytes(b"".join(struct.pack("<4I", *f) for f in felts))

    def mix_u32s(self, words):
        self.mix_bytes(struct.pack("<%dI" % len(words), *words))

    def mix_u64(self, v):
        self.mix_u32s([v & 0xFFFFFFFF, v >> 32])

    def mix_root(self, root):
        self.mix_bytes(bytes(root))

    def draw_u32s(self):
        h = hashlib.blake2s(self.digest + struct.pack("<I", self.n_draws) + b"\0").digest()
        self.n_draws += 1
        return struct.unpack("<8I", h)

    def draw_secure_felt(self):
        while True:
            w = self.draw_u32s()
            if all(x < 2 * P for x in w):
                return tuple(x % P for x in w[:4])

    def verify_pow(self, bits, nonce):
        prefix = hashlib.blake2s(struct.pack("<I", POW_PREFIX) + bytes(12) + self.digest +
                                 struct.pack("<I", bits)).digest()
        h = hashlib.blake2s(prefix + struct.pack("<Q", nonce)).digest()
        v = int.from_bytes(h[:16], "little")
        tz = 128 if v == 0 else (v & -v).bit_length() - 1
        return tz >= bits


def random_secure_point(ch):
    """randomSecureFieldPoint (circle.zig:159-167)."""
    t = ch.draw_secure_felt()
    t2 = qmul(t, t)
    inv = qinv(qadd(t2, ONE))
    return (qmul(qsub(ONE, t2), inv), qmul(qadd(t, t), inv))


# ---------------------------------------------------------------- Merkle
def _leaf(values):
    return hashlib.blake2s(LEAF_SEED + b"".join(struct.pack("<I", v) for v in values)).digest()


def _node(left, right):
    return hashlib.blake2s(NODE_SEED + left + right).digest()


def merkle_verify(root, column_log_sizes, positions, queried, witness):
    """MerkleVerifierLifted.verify (vcs_lifted/verifier.zig:93-244): queried[col][q] for the
    (possibly repeated) positions; columns hashed in (log size, index) order."""
    if not column_log_sizes:
        return
    _check(len(queried) == len(column_log_sizes), "merkle: column count")
    for col in queried:
        _check(len(col) == len(positions), "merkle: query count")
    order = sorted(range(len(positions)), key=lambda j: (positions[j], j))
    cols = sorted(range(len(column_log_sizes)), key=lambda j: (column_log_sizes[j], j))
    unique = []
    rows = {}
    for j in order:
        p = positions[j]
        row = [queried[c][j] for c in cols]
        if p in rows:
            _check(rows[p] == row, "merkle: duplicate query mismatch")
        else:
            rows[p] = row
            unique.append(p)
    layer = [(p, _leaf(rows[p])) for p in unique]
    wi = 0
    for _ in range(max(column_log_sizes)):
        nxt = []
        k = 0
        while k < len(layer):
            idx, h = layer[k]
            if k + 1 < len(layer) and layer[k + 1][0] == idx ^ 1:
                left, right = h, layer[k + 1][1]
                k += 2
            else:
                _check(wi < len(witness), "merkle: witness too short")
                w = bytes(witness[wi])
                wi += 1
                left, right = (h, w) if idx & 1 == 0 else (w, h)
                k += 1
            nxt.append((idx >> 1, _node(left, right)))
        layer = nxt
    _check(wi == len(witness), "merkle: witness too long")
    _check(len(layer) == 1 and layer[0][1] == bytes(root), "merkle: root mismatch")


# ---------------------------------------------------------------- AIRs
def _air_shape(air, st):
    """(preprocessed log sizes, main log sizes, composition log size, sampled preprocessed
    columns, mix_statement(channel), composition evaluation at the OODS point)."""
    if air == "wide_fibonacci":
        n, W = st["log_n_rows"], st["sequence_len"]
        _check(1 <= n < 31 and W >= 2, "statement out of range")

        def mix(ch):
            ch.mix_u32s([n, W])

        def comp(point, sv, alpha):
            main = sv[1]
            _check(len(main) == W and all(len(c) == 1 for c in main), "sampled value shape")
            den_inv = qinv(coset_vanishing_qm(n, point))
            acc = ZERO
            a, b = main[0][0], main[1][0]
            for col in main[2:]:
                c = col[0]
                cons = qsub(c, qadd(qmul(a, a), qmul(b, b)))
                acc = qadd(qmul(acc, alpha), qmul(cons, den_inv))
                a, b = b, c
            return acc

        return [], [n] * W, n + 1, [], mix, comp, None
    if air == "poseidon":
        li = st["log_n_instances"]
        n = li - 3
        _check(n >= 1, "statement out of range")

        def mix(ch):
            ch.mix_u32s([li])

        def comp(point, sv, alpha):
            return (li % P, POSEIDON_COLUMNS_PER_REP, POSEIDON_COLUMNS, 1)

        return [], [n] * POSEIDON_COLUMNS, n + 1, [], mix, comp, None
    if air == "xor":
        n, ls, off = st["log_size"], st["log_step"], st["offset"]
        _check(n >= 1 and ls <= n, "statement out of range")

        def mix(ch):
            ch.mix_u32s([n, ls])
            ch.mix_u64(off)

        def comp(point, sv, alpha):
            return (n, ls, off % P, 1)

        return [n, n], [n], n + 1, [0, 1], mix, comp, None
    if air == "xor_logup":
        # the XOR-lookup AIR of this backend (the pair-batched LogUp constraint of
        # frontends/riscv/air/logup.zig:93-105 over a preprocessed table component)
        n = st["log_n"]
        _check(16 <= n < 31, "statement out of range")
        claimed = ZERO  # the lookups close the bus: the claimed sum is zero

        def mix(ch):
            ch.mix_u32s([n])

        def draw(ch):
            z = ch.draw_secure_felt()
            return {"z": z, "alpha": ch.draw_secure_felt()}

        def after(ch, state):
            ch.mix_felts([claimed])

        g = point_at((-subgroup_gen(n)) % ORDER)  # minus the trace coset's step

        def prev(point):
            return qp_add(point, (qm(g[0]), qm(g[1])))

        def secure(vals):
            return qadd(qadd(vals[0], qmul(vals[1], (0, 1, 0, 0))), qadd(qmul(vals[2], (0, 0, 1, 0)),
                                                                      qmul(vals[3], (0, 0, 0, 1))))

        def comp(point, sv, alpha, state):
            pre, main, inter = sv[0], sv[1], sv[2]
            _check(len(pre) == 4 and len(main) == 4 and len(inter) == 4, "sampled value shape")
            z, al = state["z"], state["alpha"]
            al2 = qmul(al, al)

            def comb(x, y, w):
                return qsub(z, qadd(qadd(x, qmul(al, y)), qmul(al2, w)))

            isf, tx, ty, tz = (c[0] for c in pre)
            a, b, c, m = (col[0] for col in main)
            s_cur = secure([col[0] for col in inter])
            s_prev = secure([col[1] for col in inter])
            d1, d2 = comb(a, b, c), comb(tx, ty, tz)
            delta = qadd(qsub(s_cur, s_prev), qmul(isf, claimed))
            cons = qsub(qmul(qmul(delta, d1), d2), qsub(d2, qmul(m, d1)))
            return qmul(cons, qinv(coset_vanishing_qm(n, point)))

        inter = {"logs": [n] * 4, "npts": [2] * 4, "draw": draw, "after": after, "prev": prev}
        return [n] * 4, [n] * 4, n + 1, [0, 1, 2, 3], mix, comp, inter
    raise VerificationError("unknown AIR " + air)


# ---------------------------------------------------------------- FRI helpers
def _line_domain_x(init_idx, step_idx, i):
    return point_at(init_idx + step_idx * i)[0]


def _last_layer_eval(coeffs, x):
    """LinePoly.evalAtPoint (line.zig:100-120 with poly/utils.zig:14-35): coefficient index
    bit j multiplies by the (log - 1 - j)-times... folded with doublings [x, pi(x), ...]."""
    log = len(coeffs).bit_length() - 1
    doublings = []
    cur = x
    for _ in range(log):
        doublings.append(cur)
        cur = qp_double_x(cur)

    def fold(vals, fs):
        if len(vals) == 1:
            return vals[0]
        h = len(vals) // 2
        return qadd(fold(vals[:h], fs[1:]), qmul(fold(vals[h:], fs[1:]), fs[0]))

    return fold(list(coeffs), doublings)


def _rebuild(positions, evals, witness):
    """computeDecommitmentPositionsAndRebuildEvals with fold_step 1 (folding.zig:163-235):
    pairs {2k, 2k+1}, the partner taken from the witness when not queried."""
    out_pos, subsets, wi = [], [], 0
    k = 0
    while k < len(positions):
        start = (positions[k] >> 1) << 1
        sub = []
        for p in (start, start + 1):
            out_pos.append(p)
            if k < len(positions) and positions[k] == p:
                sub.append(evals[k])
                k += 1
            else:
                _check(wi < len(witness), "fri: witness too short")
                sub.append(witness[wi])
                wi += 1
        subsets.append((start, sub))
    _check(wi == len(witness), "fri: witness too long")
    return out_pos, subsets


def _coords_cols(subsets):
    cols = [[], [], [], []]
    for _, sub in subsets:
        for v in sub:
            for c in range(4):
                cols[c].append(v[c])
    return cols


def _fold_queries(positions):
    out = []
    for q in positions:
        if not out or out[-1] != q >> 1:
            out.append(q >> 1)
    return out


# ---------------------------------------------------------------- proof
def parse(proof_bytes):
    return json.loads(proof_bytes)


def _felt(v):
    _check(len(v) == 4 and all(isinstance(x, int) and 0 <= x < P for x in v), "non-canonical QM31")
    return tuple(v)


def verify_proof(proof_bytes, air, statement, pcs_config):
    """Verifies `proof_bytes` (canonical proof wire) for `air` ("wide_fibonacci", "poseidon",
    "xor") and `statement` under `pcs_config` = {pow_bits, log_blowup_factor,
    log_last_layer_degree_bound, n_queries[, fold_step, lifting_log_size]}."""
    try:
        pr = parse(proof_bytes)
    except ValueError as e:
        raise VerificationError("malformed proof: %s" % e)
    cfg = dict(pow_bits=pcs_config["pow_bits"], log_blowup_factor=pcs_config["log_blowup_factor"],
               log_last_layer_degree_bound=pcs_config["log_last_layer_degree_bound"],
               n_queries=pcs_config["n_queries"], fold_step=pcs_config.get("fold_step", 1),
               lifting_log_size=pcs_config.get("lifting_log_size"))
    pc = pr["config"]
    fc = pc["fri_config"]
    _check(pc["pow_bits"] == cfg["pow_bits"] and fc["log_blowup_factor"] == cfg["log_blowup_factor"] and
           fc["log_last_layer_degree_bound"] == cfg["log_last_layer_degree_bound"] and
           fc["n_queries"] == cfg["n_queries"] and fc["fold_step"] == cfg["fold_step"] and
           pc["lifting_log_size"] == cfg["lifting_log_size"], "proof config differs from the PCS config")
    B, n_queries, pow_bits = cfg["log_blowup_factor"], cfg["n_queries"], cfg["pow_bits"]
    llb = cfg["log_last_layer_degree_bound"]
    pre_logs, main_logs, comp_log, pre_sampled, mix_statement, comp_eval, inter = _air_shape(air, statement)
    n_trees = 4 if inter else 3

    ch = Channel()
    ch.mix_felts([(pow_bits, B, n_queries, llb)])
    if cfg["fold_step"] != 1 or cfg["lifting_log_size"] is not None:
        ch.mix_felts([(cfg["fold_step"], cfg["lifting_log_size"] or 0, 0, 0)])
    _check(cfg["fold_step"] == 1, "only fold_step 1 is supported")

    commits = pr["commitments"]
    _check(len(commits) == n_trees, "commitment count")
    for c in commits:
        _check(len(c) == 32 and all(0 <= b < 256 for b in c), "commitment shape")
    tree_logs = [[l + B for l in pre_logs], [l + B for l in main_logs]]
    ch.mix_root(commits[0])
    ch.mix_root(commits[1])
    mix_statement(ch)
    state = None
    if inter:
        # relation challenges, the interaction tree and its claimed sums (the interaction phase)
        state = inter["draw"](ch)
        ch.mix_root(commits[2])
        inter["after"](ch, state)
        tree_logs.append([l + B for l in inter["logs"]])
    _check(comp_log > 1, "composition log size")
    alpha = ch.draw_secure_felt()
    ch.mix_root(commits[-1])
    tree_logs.append([comp_log - 1 + B] * 8)
    oods = random_secure_point(ch)

    sv = [[[_felt(v) for v in col] for col in tree] for tree in pr["sampled_values"]]
    _check(len(sv) == n_trees, "sampled value trees")
    # mask points: the preprocessed columns the component samples, every main column, the
    # interaction columns at the OODS point and the previous row, and the composition columns
    # at the OODS point (components.zig:90-142, verifier.zig:109-133)
    npts = [[1 if i in pre_sampled else 0 for i in range(len(pre_logs))], [1] * len(main_logs)]
    if inter:
        npts.append(inter["npts"])
    npts.append([1] * 8)
    for t in range(n_trees):
        _check(len(sv[t]) == len(npts[t]), "sampled value columns of tree %d" % t)
        for col, k in zip(sv[t], npts[t]):
            _check(len(col) == k, "sampled values per column")
    mask_points = [oods, inter["prev"](oods) if inter else None]
    comp_cols = [col[0] for col in sv[-1]]
    left = qadd(qadd(comp_cols[0], qmul(comp_cols[1], (0, 1, 0, 0))),
                qadd(qmul(comp_cols[2], (0, 0, 1, 0)), qmul(comp_cols[3], (0, 0, 0, 1))))
    right = qadd(qadd(comp_cols[4], qmul(comp_cols[5], (0, 1, 0, 0))),
                 qadd(qmul(comp_cols[6], (0, 0, 1, 0)), qmul(comp_cols[7], (0, 0, 0, 1))))
    x = oods[0]
    for _ in range(comp_log - 2):
        x = qp_double_x(x)
    expect = comp_eval(oods, sv, alpha, state) if inter else comp_eval(oods, sv, alpha)
    _check(qadd(left, qmul(x, right)) == expect, "OODS values do not match the AIR")

    # verifyValues (pcs/verifier.zig:83-178)
    ch.mix_felts([v for tree in sv for col in tree for v in col])
    q_coeff = ch.draw_secure_felt()
    sampled_logs = [tree_logs[t][i] for t in range(n_trees) for i, k in enumerate(npts[t]) if k]
    _check(sampled_logs, "empty sampled set")
    L = max(sampled_logs)
    _check(L >= B, "lifting log size")
    fp = pr["fri_proof"]
    ch.mix_root(fp["first_layer"]["commitment"])
    first_alpha = ch.draw_secure_felt()
    bound = L - B - 1  # the column bound folded to a line
    inner = fp["inner_layers"]
    alphas = []
    for layer in inner:
        ch.mix_root(layer["commitment"])
        _check(bound - llb >= 1, "fri: too many inner layers")
        alphas.append(ch.draw_secure_felt())
        bound -= 1
    _check(bound == llb, "fri: inner layer count")
    last = [_felt(v) for v in fp["last_layer_poly"]]
    _check(len(last) <= (1 << llb) and len(last) >= 1 and (len(last) & (len(last) - 1)) == 0,
           "fri: last layer degree")
    ch.mix_felts(last)
    nonce = pr["proof_of_work"]
    _check(isinstance(nonce, int) and 0 <= nonce < (1 << 64), "pow nonce")
    _check(ch.verify_pow(pow_bits, nonce), "proof of work")
    ch.mix_u64(nonce)
    raw = []
    while len(raw) < n_queries:
        for w in ch.draw_u32s():
            if len(raw) < n_queries:
                raw.append(w & ((1 << L) - 1))
    queries = sorted(set(raw))

    # trace trees (the preprocessed tree at its own size, pcs/utils.zig:292-314)
    qv = pr["queried_values"]
    dc = pr["decommitments"]
    _check(len(qv) == n_trees and len(dc) == n_trees, "decommitment trees")
    pp_max = max(tree_logs[0]) if tree_logs[0] else 0
    for t in range(n_trees):
        logs = tree_logs[t]
        if t == 0:
            if pp_max == 0:
                pos = []
            elif L < pp_max:
                pos = [((q >> 1) << (pp_max - L + 1)) + (q & 1) for q in queries]
            else:
                pos = [((q >> (L - pp_max + 1)) << 1) + (q & 1) for q in queries]
        else:
            pos = queries
        _check(len(qv[t]) == len(logs), "queried value columns of tree %d" % t)
        vals = [[int(v) for v in col] for col in qv[t]]
        for col in vals:
            _check(all(0 <= v < P for v in col), "non-canonical queried value")
        try:
            merkle_verify(commits[t], logs, pos, vals, dc[t]["hash_witness"])
        except VerificationError as e:
            raise VerificationError("tree %d: %s" % (t, e))

    # quotient answers at the queries (fri_answers.zig:107-173): one batch per distinct point; a
    # column sampled at two points first adds its second sample at that point plus the lifting
    # domain's step doubled log-size times (samples.zig:270-285)
    batches = {}
    order = []
    pw = ONE
    flat = 0
    lift_gen = subgroup_gen(L)
    for t in range(n_trees):
        for i, col in enumerate(sv[t]):
            samples = [(mask_points[k], v) for k, v in enumerate(col)]
            if len(col) == 2:
                g = point_at((lift_gen << tree_logs[t][i]) % ORDER)
                samples.insert(0, (qp_add(mask_points[1], (qm(g[0]), qm(g[1]))), col[1]))
            for key, v in samples:
                if key not in batches:
                    batches[key] = []
                    order.append(key)
                batches[key].append((flat, v, pw))
                pw = qmul(pw, q_coeff)
            flat += 1
    qflat = [[int(v) for v in col] for tree in qv for col in tree]
    consts = []
    for key in order:
        pt = key
        cy = conj(pt[1])
        _check(cy != pt[1], "degenerate sample point")
        coeffs = []
        sa, sb = ZERO, ZERO
        for col, v, rp in batches[key]:
            a = qsub(conj(v), v)
            c = qsub(cy, pt[1])
            b = qsub(qmul(v, c), qmul(a, pt[1]))
            a, b, c = qmul(rp, a), qmul(rp, b), qmul(rp, c)
            coeffs.append((col, c))
            sa, sb = qadd(sa, a), qadd(sb, b)
        consts.append((pt, coeffs, sa, sb))
    answers = []
    for qi, pos in enumerate(queries):
        dp = point_at(circle_domain_index(L, bitrev(pos, L)))
        acc = ZERO
        for pt, coeffs, sa, sb in consts:
            num = ZERO
            for col, c in coeffs:
                num = qadd(num, qmul(qm(qflat[col][qi]), c))
            num = qsub(num, qadd(qmul(sa, qm(dp[1])), sb))
            prx, pix = (pt[0][0], pt[0][1]), (pt[0][2], pt[0][3])
            pry, piy = (pt[1][0], pt[1][1]), (pt[1][2], pt[1][3])
            d1 = _cm(((prx[0] - dp[0]) % P, prx[1]), piy)
            d2 = _cm(((pry[0] - dp[1]) % P, pry[1]), pix)
            den = ((d1[0] - d2[0]) % P, (d1[1] - d2[1]) % P)
            di = cm_inv(den)
            acc = qadd(acc, qmul(num, (di[0], di[1], 0, 0)))
        answers.append(acc)

    # FRI decommitment (fri.zig:156-214)
    fl = fp["first_layer"]
    dpos, subsets = _rebuild(queries, answers, [_felt(v) for v in fl["fri_witness"]])
    try:
        merkle_verify(fl["commitment"], [L] * 4, dpos, _coords_cols(subsets), fl["decommitment"]["hash_witness"])
    except VerificationError as e:
        raise VerificationError("fri first layer: %s" % e)
    evals = []
    for start, (f_p, f_np) in subsets:
        y = point_at(circle_domain_index(L, bitrev(start, L)))[1]
        yi = pow(y, P - 2, P)
        f0, f1 = qadd(f_p, f_np), qmul(qsub(f_p, f_np), qm(yi))
        evals.append(qadd(f0, qmul(first_alpha, f1)))
    lq = _fold_queries(queries)
    log = L - 1
    init_idx, step_idx = subgroup_gen(log + 2), subgroup_gen(log)  # LineDomain(halfOdds(log))
    for k, layer in enumerate(inner):
        dpos, subsets = _rebuild(lq, evals, [_felt(v) for v in layer["fri_witness"]])
        try:
            merkle_verify(layer["commitment"], [log] * 4, dpos, _coords_cols(subsets),
                          layer["decommitment"]["hash_witness"])
        except VerificationError as e:
            raise VerificationError("fri inner layer %d: %s" % (k, e))
        evals = []
        for start, (f_x, f_nx) in subsets:
            xx = _line_domain_x(init_idx, step_idx, bitrev(start, log))
            xi = pow(xx, P - 2, P)
            f0, f1 = qadd(f_x, f_nx), qmul(qsub(f_x, f_nx), qm(xi))
            evals.append(qadd(f0, qmul(alphas[k], f1)))
        lq = _fold_queries(lq)
        log -= 1
        init_idx, step_idx = (2 * init_idx) % ORDER, (2 * step_idx) % ORDER
    _check(log == llb + B, "fri: last layer domain")
    padded = last
    for q, e in zip(lq, evals):
        xx = _line_domain_x(init_idx, step_idx, bitrev(q, log))
        _check(_last_layer_eval(padded, qm(xx)) == e, "fri: last layer evaluation")


def verify_golden_entry(proof_bytes, entry, air):
    pc = entry["pcs_config"]
    fc = pc["fri_config"]
    cfg = dict(pow_bits=pc["pow_bits"], log_blowup_factor=fc["log_blowup_factor"],
               log_last_layer_degree_bound=fc["log_last_layer_degree_bound"], n_queries=fc["n_queries"],
               fold_step=fc["fold_step"], lifting_log_size=pc["lifting_log_size"])
    verify_proof(proof_bytes, air, entry["statement"], cfg)
