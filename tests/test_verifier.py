"""Acceptance oracle: an independent Python restatement of the reference STARK verifier
(tests/stark_verifier.py) pinned two ways, then used on the C oracle's proofs.

1. It accepts every committed golden proof of the reference (six wide-Fibonacci, two
   Poseidon, two periodic XOR proofs).
2. It rejects each proof-level mutation class of the reference's interop receipts
   (vectors/reports/interop_history/receipts/native_e2e_interop_receipt_v2/*.json,
   mutation_coverage: statement, sampled value, Merkle commitment, Merkle opening, trace
   decommitment, FRI commitment / witness / decommitment, last-layer polynomial, PoW nonce,
   PCS config n_queries on either side, fold_step, lifting_log_size), plus the inner FRI
   layers and the other two commitments.  The three metadata-policy classes of the receipts
   (prove_mode, upstream_commit, generator) are artifact fields outside the proof bytes.
3. It accepts the oracle's proofs in configurations no golden covers (secure protocol,
   seeded-uniform traces, blowup 2-3, last-layer bound 1-2, 70 queries with duplicates),
   which pins the oracle there, and the seeded-uniform trace rule is restated here.
"""
import copy
import json

import numpy as np
import pytest

import oracle
import stark_verifier as sv
from conftest import load_goldens, load_poseidon_goldens, load_xor_goldens

P = (1 << 31) - 1
WF = load_goldens()
PO = load_poseidon_goldens()
XO = load_xor_goldens()


def _cfg(e):
    pc = e["pcs_config"]
    fc = pc["fri_config"]
    return dict(pow_bits=pc["pow_bits"], log_blowup_factor=fc["log_blowup_factor"],
                log_last_layer_degree_bound=fc["log_last_layer_degree_bound"], n_queries=fc["n_queries"],
                fold_step=fc["fold_step"], lifting_log_size=pc["lifting_log_size"])


GOLDEN_CASES = [(g, "wide_fibonacci") for g in WF] + [(g, "poseidon") for g in PO] + [(g, "xor") for g in XO]


@pytest.mark.parametrize("g,air", GOLDEN_CASES, ids=[g["name"] for g, _ in GOLDEN_CASES])
def test_verifier_accepts_golden(g, air):
    sv.verify_proof(g["proof"], air, g["statement"], _cfg(g))


def _bump(v):
    return (v + 1) % P


def _flip(h):
    h = list(h)
    h[0] ^= 1
    return h


def _mutations(pr):
    """(class, mutated proof dict | None, statement delta, config delta)."""
    out = []

    def m(name, fn, st=None, cf=None):
        q = copy.deepcopy(pr)
        if fn is not None:
            fn(q)
        out.append((name, q, st or {}, cf or {}))

    def first_nonempty(trees, key):
        for t, tree in enumerate(trees):
            if tree[key] if key else tree:
                return t
        raise AssertionError("no non-empty tree")

    m("statement", None, st="bump")
    t_sv = first_nonempty(pr["sampled_values"], None)
    m("transcript_bound_sampled_value", lambda q: q["sampled_values"][t_sv][0][0].__setitem__(0, _bump(
        q["sampled_values"][t_sv][0][0][0])))
    for t in range(3):
        m("merkle_commitment[%d]" % t, lambda q, t=t: q["commitments"].__setitem__(t, _flip(q["commitments"][t])))
    t_qv = first_nonempty(pr["queried_values"], None)
    m("merkle_opening", lambda q: q["queried_values"][t_qv][0].__setitem__(0, _bump(q["queried_values"][t_qv][0][0])))
    t_dc = first_nonempty(pr["decommitments"], "hash_witness")
    m("merkle_trace_decommitment", lambda q: q["decommitments"][t_dc]["hash_witness"].__setitem__(
        0, _flip(q["decommitments"][t_dc]["hash_witness"][0])))
    fl = "fri_proof"
    m("fri_commitment", lambda q: q[fl]["first_layer"].__setitem__("commitment", _flip(q[fl]["first_layer"]["commitment"])))
    if pr[fl]["first_layer"]["fri_witness"]:
        m("fri_witness", lambda q: q[fl]["first_layer"]["fri_witness"][0].__setitem__(
            0, _bump(q[fl]["first_layer"]["fri_witness"][0][0])))
    m("fri_decommitment", lambda q: q[fl]["first_layer"]["decommitment"]["hash_witness"].__setitem__(
        0, _flip(q[fl]["first_layer"]["decommitment"]["hash_witness"][0])))
    m("fri_inner_commitment", lambda q: q[fl]["inner_layers"][0].__setitem__(
        "commitment", _flip(q[fl]["inner_layers"][0]["commitment"])))
    k_in = max(k for k, layer in enumerate(pr[fl]["inner_layers"]) if layer["decommitment"]["hash_witness"])
    m("fri_inner_decommitment", lambda q: q[fl]["inner_layers"][k_in]["decommitment"]["hash_witness"].__setitem__(
        0, _flip(q[fl]["inner_layers"][k_in]["decommitment"]["hash_witness"][0])))
    m("fri_last_layer_polynomial", lambda q: q[fl]["last_layer_poly"][0].__setitem__(
        0, _bump(q[fl]["last_layer_poly"][0][0])))
    m("pow_nonce", lambda q: q.__setitem__("proof_of_work", q["proof_of_work"] + 1))
    m("artifact_pcs_config", None, cf={"n_queries": pr["config"]["fri_config"]["n_queries"] + 1})
    m("proof_pcs_config", lambda q: q["config"]["fri_config"].__setitem__("n_queries", q["config"]["fri_config"]["n_queries"] + 1))
    m("outer_fold_step", None, cf={"fold_step": 2})
    m("outer_lifting_log_size", None, cf={"lifting_log_size": 5})
    return out


def _bump_statement(air, st):
    st = dict(st)
    key = {"wide_fibonacci": "sequence_len", "poseidon": "log_n_instances", "xor": "offset"}[air]
    st[key] += 1
    return st


MUTATION_BASES = [(WF[1], "wide_fibonacci"), (WF[0], "wide_fibonacci"), (PO[0], "poseidon"), (XO[0], "xor")]


@pytest.mark.parametrize("g,air", MUTATION_BASES, ids=[g["name"] for g, _ in MUTATION_BASES])
def test_verifier_rejects_every_mutation_class(g, air):
    pr = json.loads(g["proof"])
    muts = _mutations(pr)
    assert len(muts) >= 17
    for name, q, st, cf in muts:
        stmt = _bump_statement(air, g["statement"]) if st == "bump" else g["statement"]
        cfg = dict(_cfg(g), **cf)
        data = json.dumps(q, separators=(",", ":")).encode()
        with pytest.raises(sv.VerificationError):
            sv.verify_proof(data, air, stmt, cfg)
            pytest.fail("mutation %s was accepted" % name)
    # the unmutated re-serialisation is the golden byte string and is accepted
    assert json.dumps(pr, separators=(",", ":")).encode() == g["proof"]


# ---------------------------------------------------------------- oracle proofs no golden covers
ORACLE_CASES = [
    # (air, statement, pcs config, trace mode/seed)
    ("wide_fibonacci", dict(log_n_rows=12, sequence_len=16), dict(pow_bits=26, log_blowup_factor=1,
                                                                  log_last_layer_degree_bound=0, n_queries=70), (1, 7)),
    ("wide_fibonacci", dict(log_n_rows=12, sequence_len=16), dict(pow_bits=26, log_blowup_factor=1,
                                                                  log_last_layer_degree_bound=0, n_queries=70), (0, 0)),
    ("wide_fibonacci", dict(log_n_rows=6, sequence_len=8), dict(pow_bits=4, log_blowup_factor=2,
                                                                log_last_layer_degree_bound=1, n_queries=70), (1, 3)),
    ("wide_fibonacci", dict(log_n_rows=7, sequence_len=5), dict(pow_bits=0, log_blowup_factor=3,
                                                                log_last_layer_degree_bound=2, n_queries=20), (1, 11)),
    ("wide_fibonacci", dict(log_n_rows=3, sequence_len=4), dict(pow_bits=2, log_blowup_factor=1,
                                                                log_last_layer_degree_bound=0, n_queries=70), (1, 5)),
    ("wide_fibonacci", dict(log_n_rows=16, sequence_len=64), dict(pow_bits=12, log_blowup_factor=1,
                                                                  log_last_layer_degree_bound=2, n_queries=16), (1, 99)),
    ("poseidon", dict(log_n_instances=8), dict(pow_bits=6, log_blowup_factor=2, log_last_layer_degree_bound=1,
                                               n_queries=12), None),
    ("xor", dict(log_size=9, log_step=3, offset=5), dict(pow_bits=5, log_blowup_factor=3,
                                                         log_last_layer_degree_bound=2, n_queries=30), None),
    ("xor", dict(log_size=12, log_step=0, offset=(1 << 40) + 17), dict(pow_bits=8, log_blowup_factor=1,
                                                                        log_last_layer_degree_bound=0, n_queries=70), None),
]


def oracle_proof(air, st, cfg, trace):
    if air == "wide_fibonacci":
        mode, seed = trace
        return oracle.prove_wide_fibonacci(st["log_n_rows"], st["sequence_len"], trace_mode=mode, trace_seed=seed, **cfg)
    if air == "poseidon":
        return oracle.prove_poseidon(st["log_n_instances"], **cfg)
    return oracle.prove_xor(st["log_size"], st["log_step"], st["offset"], **cfg)


@pytest.mark.parametrize("case", ORACLE_CASES, ids=["%s-%s-%s" % (c[0], "-".join(map(str, c[1].values())),
                                                                    "-".join(map(str, c[2].values())))
                                                    for c in ORACLE_CASES])
def test_verifier_accepts_oracle_proof(case):
    air, st, cfg, trace = case
    proof = oracle_proof(air, st, cfg, trace)
    sv.verify_proof(proof, air, st, cfg)
    # and rejects it for a neighbouring statement
    with pytest.raises(sv.VerificationError):
        sv.verify_proof(proof, air, _bump_statement(air, st), cfg)


# ---------------------------------------------------------------- seeded-uniform trace rule
def splitmix64(x):
    """SplitMix64 output function (Steele, Lea, Flood 2014; Vigna's splitmix64.c)."""
    M = (1 << 64) - 1
    z = (x + 0x9E3779B97F4A7C15) & M
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
    return z ^ (z >> 31)


def test_splitmix64_known_answers():
    # the first outputs of the published generator seeded with state 0 (state += golden gamma
    # before each output): 0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4
    assert splitmix64(0) == 0xE220A8397B1DCDAF
    assert splitmix64(0x9E3779B97F4A7C15) == 0x6E789E6AA1B965F4


@pytest.mark.parametrize("log_n,seed", [(3, 0), (5, 42), (9, (1 << 40) + 3)])
def test_seeded_uniform_trace_rule(log_n, seed):
    # mode 1: first two columns of row s = splitmix64(seed ^ k * 0xD1B54A32D192ED03) mod p for
    # k = 2s, 2s + 1; the rest follows the recurrence c = a^2 + b^2
    M = (1 << 64) - 1
    n = 1 << log_n
    t = oracle.widefib_trace(log_n, 4, mode=1, seed=seed)
    for s in range(n):
        a = splitmix64(seed ^ ((2 * s * 0xD1B54A32D192ED03) & M)) % P
        b = splitmix64(seed ^ (((2 * s + 1) * 0xD1B54A32D192ED03) & M)) % P
        assert (int(t[0][s]), int(t[1][s])) == (a, b)
        assert int(t[2][s]) == (a * a + b * b) % P
        assert int(t[3][s]) == (b * b + int(t[2][s]) ** 2) % P


# ---------------------------------------------------------------- XOR-lookup AIR (LogUp proof)
LOGUP_CASES = [
    (16, 7, dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)),
    (16, 99, dict(pow_bits=4, log_blowup_factor=2, log_last_layer_degree_bound=1, n_queries=20)),
    (17, 3, dict(pow_bits=0, log_blowup_factor=1, log_last_layer_degree_bound=2, n_queries=40)),
]


@pytest.mark.parametrize("log_n,seed,cfg", LOGUP_CASES, ids=["16-functional", "16-blowup2", "17-last2"])
def test_verifier_accepts_logup_xor_proof(log_n, seed, cfg):
    """The XOR-lookup AIR's whole proof (is_first and table tuple preprocessed; a, b, c and the
    multiplicities committed; the pair-batched cumulative column as the interaction tree,
    sampled at the OODS point and the previous row): the composition check at the OODS point
    ties the committed composition to the pair constraint, so acceptance pins the oracle's
    interaction column, its previous-row mask and the two-point quotient batches."""
    proof = oracle.prove_logup_xor(log_n, seed, **cfg)
    sv.verify_proof(proof, "xor_logup", {"log_n": log_n}, cfg)
    assert len(json.loads(proof)["commitments"]) == 4


def test_verifier_rejects_logup_xor_mutations():
    cfg = LOGUP_CASES[0][2]
    base = json.loads(oracle.prove_logup_xor(16, 7, **cfg))

    def rejected(mutate, statement=None):
        pr = copy.deepcopy(base)
        mutate(pr)
        with pytest.raises(sv.VerificationError):
            sv.verify_proof(json.dumps(pr, separators=(",", ":")).encode(), "xor_logup",
                            statement or {"log_n": 16}, cfg)

    def bump(v):
        v[0] = (v[0] + 1) % P

    rejected(lambda pr: None, {"log_n": 17})                                   # statement
    rejected(lambda pr: bump(pr["sampled_values"][2][1][1]))                   # previous-row sample
    rejected(lambda pr: bump(pr["sampled_values"][2][0][0]))                   # current interaction sample
    rejected(lambda pr: bump(pr["sampled_values"][1][3][0]))                   # multiplicity sample
    rejected(lambda pr: pr["commitments"][2].__setitem__(0, pr["commitments"][2][0] ^ 1))  # interaction root
    rejected(lambda pr: pr["queried_values"][2][1].__setitem__(0, (pr["queried_values"][2][1][0] + 1) % P))
    rejected(lambda pr: pr["sampled_values"][2][0].pop())                      # mask shape
