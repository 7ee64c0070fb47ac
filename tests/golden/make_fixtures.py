"""Extract the reference's committed wide-Fibonacci, Poseidon and XOR golden proofs into small fixtures.

Reads only data files the reference holds (proof artifacts written by its CPU
backend and cross-verified against pinned Rust Stwo):
  vectors/reports/interop_history/objects/sha256/ac/ac89...json      (5x16, pow 0)
  vectors/reports/benchmark_history/runs/*/bundle/tree/row-*wide*/cpu.proof-artifact.json
  conformance/evidence/metrics-v2/w1-log20x100/proof-artifact.json   (20x100)
  vectors/reports/benchmark_history/runs/*/bundle/tree/row-*-{poseidon,xor}-*/cpu.proof-artifact.json
Each fixture is the canonical proof byte string (compact JSON wire) plus its
statement and PCS config.  Run from the repository root:
    python tests/golden/make_fixtures.py /root/reference
"""
import glob
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    # Interop store: only objects a receipt lists as an accepted proof (the store also
    # holds deliberately tampered negative cases).
    paths = []
    for r in sorted(glob.glob(os.path.join(ref, "vectors/reports/interop_history/receipts/*/*.json"))):
        for a in json.load(open(r)).get("artifacts", []):
            if a.get("example") == "wide_fibonacci" and a.get("role") == "accepted_proof":
                paths.append(os.path.join(ref, "vectors/reports/interop_history", a["object_path"]))
    paths += glob.glob(os.path.join(ref, "vectors/reports/benchmark_history/runs/*/bundle/tree/*wide_fibonacci*/cpu.proof-artifact.json"))
    paths += [os.path.join(ref, "conformance/evidence/metrics-v2/w1-log20x100/proof-artifact.json")]
    index = {}
    for p in sorted(paths):
        try:
            art = json.load(open(p))
        except Exception:
            continue
        if art.get("example") != "wide_fibonacci" or "proof_bytes_hex" not in art:
            continue
        if art.get("prove_mode") not in (None, "prove"):
            continue
        proof = bytes.fromhex(art["proof_bytes_hex"])
        # Keep only proofs that parse and whose embedded config equals the artifact config
        # (the interop store also holds deliberately mutated negative cases).
        try:
            decoded = json.loads(proof)
        except Exception:
            continue
        if decoded.get("config") != art["pcs_config"]:
            continue
        st = art["wide_fibonacci_statement"]
        name = "wide_fibonacci_log%d_seq%d_pow%d" % (st["log_n_rows"], st["sequence_len"], art["pcs_config"]["pow_bits"])
        sha = hashlib.sha256(proof).hexdigest()
        if name in index:
            if index[name]["sha256"] != sha:
                index[name].setdefault("conflicts", []).append(os.path.relpath(p, ref))
            continue
        fname = name + ".proof.json"
        with open(os.path.join(HERE, fname), "wb") as f:
            f.write(proof)
        index[name] = {
            "file": fname,
            "statement": st,
            "pcs_config": art["pcs_config"],
            "sha256": sha,
            "bytes": len(proof),
            "source": os.path.relpath(p, ref),
        }
    with open(os.path.join(HERE, "wide_fibonacci_goldens.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    for k, v in sorted(index.items()):
        print(k, v["bytes"], v["sha256"][:16], v.get("conflicts", ""))
    poseidon(ref)


def poseidon(ref):
    example_rows(ref, "poseidon", lambda st: "poseidon_log%d" % st["log_n_instances"])
    example_rows(ref, "xor", lambda st: "xor_log%d_step%d_offset%d" % (st["log_size"], st["log_step"], st["offset"]))


def example_rows(ref, example, stem):
    """Benchmark-row proofs of one example AIR into <example>_goldens.json."""
    index = {}
    for p in sorted(glob.glob(os.path.join(ref, "vectors/reports/benchmark_history/runs/*/bundle/tree/*-%s-*/cpu.proof-artifact.json" % example))):
        art = json.load(open(p))
        if art.get("example") != example or "proof_bytes_hex" not in art or art.get("prove_mode") not in (None, "prove"):
            continue
        proof = bytes.fromhex(art["proof_bytes_hex"])
        if json.loads(proof).get("config") != art["pcs_config"]:
            continue
        st = art[example + "_statement"]
        name = stem(st) + "_pow%d" % art["pcs_config"]["pow_bits"]
        sha = hashlib.sha256(proof).hexdigest()
        if name in index:
            if index[name]["sha256"] != sha:
                index[name].setdefault("conflicts", []).append(os.path.relpath(p, ref))
            continue
        fname = name + ".proof.json"
        with open(os.path.join(HERE, fname), "wb") as f:
            f.write(proof)
        index[name] = {"file": fname, "statement": st, "pcs_config": art["pcs_config"], "sha256": sha,
                       "bytes": len(proof), "source": os.path.relpath(p, ref)}
    with open(os.path.join(HERE, example + "_goldens.json"), "w") as f:
        json.dump(index, f, indent=1, sort_keys=True)
    for k, v in sorted(index.items()):
        print(k, v["bytes"], v["sha256"][:16], v.get("conflicts", ""))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
