"""External verifier command for the prover session's tests (test infrastructure).

    python verify_cli.py --statement STATEMENT.json --proof PROOF

STATEMENT.json is the session's statement file: {"air": ..., "statement": {...},
"pcs_config": {pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries}}.
Exit status 0 when stark_verifier.verify_proof accepts the proof, 1 when it rejects
it (the reason on stderr), 2 on bad arguments.  The session runs it once per proof,
as the reference's session runs its external verifier
(src/tools/metal_prover_session/verification.zig).
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

import stark_verifier as sv  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--statement", required=True)
    ap.add_argument("--proof", required=True)
    a = ap.parse_args(argv)
    st = json.load(open(a.statement))
    proof = open(a.proof, "rb").read()
    try:
        sv.verify_proof(proof, st["air"], st["statement"], st["pcs_config"])
    except sv.VerificationError as e:
        print("rejected: %s" % e, file=sys.stderr)
        return 1
    print("accepted %s %d bytes" % (st["air"], len(proof)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
