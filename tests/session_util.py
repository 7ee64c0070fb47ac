"""Helpers that drive the stwo_rocm_session executable over its JSONL protocol (test infrastructure)."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SESSION = os.path.join(ROOT, "stwo-zig_amd", "lib", "stwo_rocm_session")
VERIFY_CLI = os.path.join(ROOT, "tests", "verify_cli.py")
PROTOCOL = "stwo-zig-rocm-prover-session"
VERSION = 1
FIXTURE = os.path.join(ROOT, "tests", "golden", "session_reference_reports.json")


def prove_frame(seq, workload, outdir, proof_protocol="functional", warmups=0, samples=1, budget="16",
                request_id=None, **override):
    f = dict(protocol=PROTOCOL, version=VERSION, type="prove", sequence=seq,
             request_id=request_id or "req-%d" % seq, workload=workload, proof_protocol=proof_protocol,
             warmups=warmups, samples=samples,
             outputs=dict(proof=os.path.join(outdir, "proof-%d.json" % seq),
                          report=os.path.join(outdir, "report-%d.json" % seq)),
             budget_gib=budget)
    f.update(override)
    return f


def shutdown_frame(next_sequence):
    return dict(protocol=PROTOCOL, version=VERSION, type="shutdown", next_sequence=next_sequence)


def run_session(frames, args=(), timeout=600, raw_lines=None):
    """Feeds the frames (dicts, or raw strings) and returns (exit status, answered frames, stderr)."""
    if raw_lines is None:
        raw_lines = [f if isinstance(f, str) else json.dumps(f, separators=(",", ":")) for f in frames]
    data = "".join(line + "\n" for line in raw_lines)
    r = subprocess.run([SESSION] + list(args), input=data, capture_output=True, text=True, timeout=timeout)
    out = [json.loads(line) for line in r.stdout.splitlines() if line.strip()]
    return r.returncode, out, r.stderr


def verifier_args():
    return ["--verifier", sys.executable, "--verifier-arg", VERIFY_CLI]


def descriptor(workload, protocol):
    """The workload descriptor the session hashes (native/examples.zig:125-175 layout)."""
    name = workload["example"]
    keys = dict(wide_fibonacci=("log_n_rows", "sequence_len"), poseidon=("log_n_instances",),
                xor=("log_size", "log_step", "offset"), xor_logup=("log_n", "seed"))[name]
    s = "native-proof-workload-v3|example=%s" % name + "".join("|%s=%d" % (k, workload[k]) for k in keys)
    return s + ("|protocol=%s|pow_bits=%d|log_blowup_factor=%d|log_last_layer_degree_bound=%d|n_queries=%d"
                "|fold_step=%d" % (protocol["name"], protocol["pow_bits"], protocol["log_blowup_factor"],
                                   protocol["log_last_layer_degree_bound"], protocol["n_queries"],
                                   protocol["fold_step"]))


def load_fixture():
    return json.load(open(FIXTURE))
