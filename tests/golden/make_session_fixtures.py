"""Extract what the reference's own schema-7 reports fix about a proof request into a small fixture.

Reads only data files the reference holds: the native runner's reports of one benchmark run
(vectors/reports/benchmark_history/runs/<run>/bundle/tree/row-*/metal.stdout.json and
cpu.stdout.json, schema_version 7).  For each wide-Fibonacci, Poseidon and periodic XOR row it
keeps the workload, its parameters, the proof protocol, the workload descriptor digest and the
canonical proof's byte count and SHA-256; and it keeps the key layout (top level and one level
down) of one report, so the session's reports can be checked against both.  Run from the
repository root:
    python tests/golden/make_session_fixtures.py /root/reference
"""
import glob
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SUPPORTED = ("wide_fibonacci", "poseidon", "xor")


def main(ref):
    rows = {}
    layout = None
    layout_src = None
    for f in sorted(glob.glob(os.path.join(ref, "vectors/reports/benchmark_history/runs/*/bundle/tree/*/*.stdout.json"))):
        d = json.load(open(f))
        if d.get("schema_version") != 7:
            continue
        w, p = d["workload"], d["protocol"]
        if w["name"] not in SUPPORTED:
            continue
        key = (w["name"], json.dumps(w["parameters"], sort_keys=True), json.dumps(p, sort_keys=True))
        s0 = d["proof"]["samples"][0]
        e = dict(workload=w["name"], parameters=w["parameters"], protocol=p,
                 descriptor_sha256=w["descriptor_sha256"], proof_bytes=s0["bytes"], proof_sha256=s0["sha256"],
                 trace_log_rows=w["trace_log_rows"], committed_columns=w["committed_columns"],
                 native_unit=w["native_unit"], native_units=w["native_units"],
                 source=os.path.relpath(f, ref))
        prev = rows.get(key)
        if prev:
            for k in ("descriptor_sha256", "proof_bytes", "proof_sha256"):
                assert prev[k] == e[k], (f, k)
        else:
            rows[key] = e
        if layout is None and d.get("backend_telemetry") is not None:
            layout = {k: (sorted(v.keys()) if isinstance(v, dict) else None) for k, v in d.items()}
            layout_src = os.path.relpath(f, ref)
    out = dict(workloads=sorted(rows.values(), key=lambda e: (e["workload"], json.dumps(e["parameters"]))),
               report_layout=layout, report_layout_source=layout_src)
    path = os.path.join(HERE, "session_reference_reports.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1, sort_keys=True)
        fh.write("\n")
    print("wrote", path, len(rows), "workloads")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
