"""The reference's generic proving transaction for the wide-Fibonacci AIR, driven only through
the resident hooks of the C-ABI, in the order the generic prover calls them on a backend B:

  examples/wide_fibonacci.zig:181-343 -> examples/common/prover_transaction.zig:222-339
  -> prover/prove.zig:209-389 (composition, OODS) -> prover/pcs/scheme.zig:571-800
  (sampled values, quotients, FRI, PoW, decommit) -> prover/fri.zig:163-621
  -> interop/proof_wire.zig:15-76 (the wire).

Each step names the hook a Zig RocmBackend would forward to the C-ABI:
  PcsConfig.mixInto, mixRoot, mixU32s, drawSecureFelt, drawU32s -> stwo_rocm_channel_*
  fillQuadraticRecurrenceTrace                                  -> stwo_rocm_widefib_trace
  prepareAndCommitOwned / prepareAndCommitPolys                 -> stwo_rocm_prepare_and_commit_*
  computeCompositionEvaluation                                  -> stwo_rocm_widefib_composition
  interpolateCircleBuffers / evaluateCircleBuffers              -> stwo_rocm_circle_*
  evaluateCoefficientPlans                                      -> stwo_rocm_eval_at_point
  commitLazyMerkle                                              -> stwo_rocm_compute_quotients (+ tree)
  foldCircleIntoLine                                            -> stwo_rocm_fri_fold_circle_into_line
  commitFriLayers                                               -> stwo_rocm_fri_commit_layers
  grind                                                         -> stwo_rocm_pow_grind
  MerkleTree.decommit + queried values                          -> stwo_rocm_tree_decommit, stwo_rocm_gather_values

The host keeps only what the reference's host keeps: the transcript order, the OODS point from
the drawn felt, the last-layer interpolation of a few values and the JSON wire.  Test code: the
proof it returns is compared with the golden proofs and with the monolithic entry point.
"""
import json

import numpy as np

import stark_verifier as sv

P = sv.P


def _half_odds(log):
    """Coset.halfOdds(log): (initial index, step index) (core/circle.zig)."""
    return 1 << (31 - log - 2), 1 << (31 - log)


def _bitrev_list(v):
    n = len(v)
    lg = n.bit_length() - 1
    return [v[sv.bitrev(i, lg)] for i in range(n)]


def last_layer_poly(values, log_size, log_degree_bound):
    """commitLastLayer (fri.zig:594-621): LineEvaluation.interpolate over LineDomain(halfOdds(log))
    (line_evaluation.zig:97-151), degree check, LinePoly.fromOrderedCoefficients (line.zig:134-142)."""
    v = _bitrev_list([tuple(int(x) for x in q) for q in values])
    n = len(v)
    init, step = _half_odds(log_size)
    lg = log_size
    while lg > 0:
        csz, half = 1 << lg, 1 << (lg - 1)
        for s0 in range(0, n, csz):
            for i in range(half):
                x = sv.point_at(init + step * i)[0]
                ix = pow(x, P - 2, P)
                v0, v1 = v[s0 + i], v[s0 + half + i]
                v[s0 + i] = sv.qadd(v0, v1)
                v[s0 + half + i] = sv.qmul(sv.qsub(v0, v1), sv.qm(ix))
        init, step, lg = 2 * init, 2 * step, lg - 1
    ni = sv.qm(pow(n % P, P - 2, P))
    coeffs = _bitrev_list([sv.qmul(x, ni) for x in v])
    bound = 1 << log_degree_bound
    if any(c != sv.ZERO for c in coeffs[bound:]):
        raise ValueError("InvalidLastLayerDegree")
    return _bitrev_list(coeffs[:bound])


def _fri_layer_wire(ctx, tree, col_ptrs, log_size, words_per_value, positions):
    """decommit of one FRI layer with fold_step 1 (prover/fri.zig:626-740): the pairs {2k, 2k+1}
    of the queried positions, the partners that are not queried as the FRI witness, and the
    Merkle decommitment at the pairs."""
    dpos, partners = [], []
    i = 0
    while i < len(positions):
        key = positions[i] >> 1
        j = i
        while j < len(positions) and positions[j] >> 1 == key:
            j += 1
        at = i
        for p in (2 * key, 2 * key + 1):
            dpos.append(p)
            if at < j and positions[at] == p:
                at += 1
            else:
                partners.append(p)
        i = j
    if words_per_value == 1:  # the 4 coordinate columns of the first layer
        v = ctx.gather_values(col_ptrs, [log_size] * 4, log_size, partners)
        wit = [[int(v[k, q, 0]) for k in range(4)] for q in range(len(partners))]
    else:  # the AoS line of an inner layer
        v = ctx.gather_values(col_ptrs, [log_size], log_size, partners, words_per_value=4)
        wit = [[int(x) for x in v[0, q]] for q in range(len(partners))]
    return {"fri_witness": wit,
            "decommitment": {"hash_witness": [list(h) for h in tree.decommit_witness(dpos)]},
            "commitment": list(tree.root)}


def prove_wide_fibonacci(ctx, log_n_rows, sequence_len, pow_bits=10, log_blowup_factor=1,
                         log_last_layer_degree_bound=0, n_queries=3, trace_mode=0, trace_seed=0):
    n, W, B = log_n_rows, sequence_len, log_blowup_factor
    N, L = 1 << n, n + B
    owned = []  # device buffers and trees released at the end (ResidentStorage destroyFn)

    def buf(nbytes):
        b = ctx.alloc(nbytes)
        owned.append(b)
        return b

    ch = ctx.channel()
    try:
        # PcsConfig.mixInto (core/pcs/mod.zig:32-52)
        ch.mix_felts([[pow_bits, B, n_queries, log_last_layer_degree_bound]])
        # tree 0: the preprocessed tree, no columns for this AIR
        t0 = ctx.prepare_and_commit_owned([], [], B, [])
        owned.append(t0)
        ch.mix_root(t0.root)
        # tree 1: trace -> coefficients (in place) -> LDE -> Merkle
        trace = buf(4 * N * W)
        ctx.fill_quadratic_recurrence_trace(n, W, trace, mode=trace_mode, seed=trace_seed)
        coeffs = [trace.offset(i * N) for i in range(W)]
        lde = buf(4 * (N << B) * W)
        lde_cols = [lde.offset(i * (N << B)) for i in range(W)]
        t1 = ctx.prepare_and_commit_owned(coeffs, [n] * W, B, lde_cols)
        owned.append(t1)
        ch.mix_root(t1.root)
        ch.mix_u32s([n, W])  # mixStatement (wide_fibonacci.zig:496-498)
        # composition (prove.zig:237-318) on CanonicCoset(n + 1)
        alpha = ch.draw_secure_felt()
        CN = 2 * N
        if B == 1:
            ev_cols = lde_cols
        else:
            ev = buf(4 * CN * W)
            ev_cols = [ev.offset(i * CN) for i in range(W)]
            ctx.evaluate_circle_buffers(coeffs, n, 1, ev_cols)
        comp = buf(16 * CN)
        comp_cols = [comp.offset(k * CN) for k in range(4)]
        ctx.compute_composition_evaluation(ev_cols, n, alpha, comp_cols)
        ctx.interpolate_circle_buffers(comp_cols, n + 1)
        # splitAtMid into 8 coefficient columns of 2^n, then commitPolys
        halves = [comp.offset(k * CN) for k in range(4)] + [comp.offset(k * CN + N) for k in range(4)]
        lde2 = buf(4 * (N << B) * 8)
        lde2_cols = [lde2.offset(i * (N << B)) for i in range(8)]
        t2 = ctx.prepare_and_commit_polys(halves, [n] * 8, B, lde2_cols)
        owned.append(t2)
        ch.mix_root(t2.root)
        # OODS point (circle.zig:159-167) and sampled values (sampled_values.zig:67-80): every
        # column lives on the lifting domain 2^L, so the point is used as drawn
        t = tuple(ch.draw_secure_felt())
        t2_ = sv.qmul(t, t)
        inv = sv.qinv(sv.qadd(t2_, sv.ONE))
        oods = (sv.qmul(sv.qsub(sv.ONE, t2_), inv), sv.qmul(sv.qadd(t, t), inv))
        point8 = list(oods[0]) + list(oods[1])
        sv1 = ctx.eval_at_point(coeffs, n, point8)
        sv2 = ctx.eval_at_point(halves, n, point8)
        values = [[int(x) for x in r] for r in np.concatenate([sv1, sv2])]
        ch.mix_felts(values)
        # quotients + first FRI layer (commitLazyMerkle, fri.zig:335-370)
        qalpha = ch.draw_secure_felt()
        LN = 1 << L
        quot = buf(16 * LN)
        qc = [quot.offset(k * LN) for k in range(4)]
        qcols = lde_cols + lde2_cols
        nq = len(qcols)
        tq = ctx.compute_quotients(L, qcols, [L] * nq, list(range(nq)), [point8] * nq, values, qalpha, qc,
                                   commit=True)
        owned.append(tq)
        ch.mix_root(tq.root)
        # fold the circle into the first line (fri.zig:393-426), then commitFriLayers
        fa = ch.draw_secure_felt()
        line = buf(16 * (LN // 2))
        line.upload(np.zeros(4 * (LN // 2), dtype=np.uint32))
        ctx.fold_circle_into_line(qc, L, fa, line)
        last_log = log_last_layer_degree_bound + B
        inner = []  # (tree, line buffer, log)
        if L - 1 > last_log:
            below = [buf(16 << (L - 2 - k)) for k in range(L - 2 - last_log)]
            last = buf(16 << last_log)
            trees = ctx.commit_fri_layers(ch, line, L - 1, last_log, below, last)
            owned.extend(trees)
            lines = [line] + below
            inner = [(trees[k], lines[k], L - 1 - k) for k in range(len(trees))]
        else:
            last = line
        lastv = last.download(count=4 << last_log).reshape(-1, 4)
        lastp = last_layer_poly(lastv, last_log, log_last_layer_degree_bound)
        ch.mix_felts(lastp)
        # proof of work (scheme.zig:733-735)
        digest, _ = ch.state()
        nonce = ctx.grind(digest, pow_bits)
        ch.mix_u64(nonce)
        # queries (core/queries.zig:8-60)
        words = ch.draw_u32s((n_queries + 7) // 8)[:n_queries]
        qs = sorted(set(w & (LN - 1) for w in words))
        # decommitments and queried values (prover/vcs_lifted/decommit.zig)
        decs, qvals = [], []
        for tree, cols in ((t0, []), (t1, lde_cols), (t2, lde2_cols)):
            decs.append({"hash_witness": [list(h) for h in tree.decommit_witness(qs)] if cols else []})
            v = ctx.gather_values(cols, [L] * len(cols), L, qs) if cols else np.zeros((0, 0, 1))
            qvals.append([[int(x) for x in v[i, :, 0]] for i in range(len(cols))])
        first = _fri_layer_wire(ctx, tq, qc, L, 1, qs)
        inner_wire = []
        lq = qs
        for tree, ln, lg in inner:
            lq = sorted(set(q >> 1 for q in lq))
            inner_wire.append(_fri_layer_wire(ctx, tree, [ln], lg, 4, lq))
        proof = {
            "config": {"pow_bits": pow_bits,
                       "fri_config": {"log_blowup_factor": B, "log_last_layer_degree_bound": log_last_layer_degree_bound,
                                      "n_queries": n_queries, "fold_step": 1},
                       "lifting_log_size": None},
            "commitments": [list(t0.root), list(t1.root), list(t2.root)],
            "sampled_values": [[], [[v] for v in values[:W]], [[v] for v in values[W:]]],
            "decommitments": decs,
            "queried_values": qvals,
            "proof_of_work": nonce,
            "fri_proof": {"first_layer": first, "inner_layers": inner_wire,
                          "last_layer_poly": [list(c) for c in lastp]},
        }
        return json.dumps(proof, separators=(",", ":")).encode()
    finally:
        ch.close()
        for o in reversed(owned):
            o.free()
