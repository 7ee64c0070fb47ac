"""The session front door on the device: one persistent stwo_rocm_session process proves a
sequence of requests over its JSONL protocol, every proof checked by the external verifier
command before it is published, and writes a schema-7 report per request.

Pinned against the reference's own data: the published proofs of the wide-Fibonacci, Poseidon and
periodic XOR requests are the reference's golden proofs, byte for byte, and carry the proof
digests and workload descriptor digests of the reference's schema-7 reports
(vectors/reports/benchmark_history, tests/golden/session_reference_reports.json); the
XOR-lookup proof equals the oracle's.  The report's summaries are recomputed from its samples
(median and median absolute deviation, src/prover/native/statistics.zig:10-36) and its key
layout is compared with the reference report's."""
import hashlib
import json
import os
import statistics

import pytest

import oracle
from conftest import load_goldens, load_poseidon_goldens, load_xor_goldens
from session_util import load_fixture, prove_frame, run_session, shutdown_frame, verifier_args

pytestmark = pytest.mark.gpu

FUNC = dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3)
REQUESTS = [
    (dict(example="wide_fibonacci", log_n_rows=10, sequence_len=8), "functional", 1, 3),
    (dict(example="poseidon", log_n_instances=10), "functional", 0, 1),
    (dict(example="xor", log_size=14, log_step=2, offset=3), "functional", 0, 2),
    (dict(example="xor_logup", log_n=16, seed=7), "functional", 0, 1),
    (dict(example="wide_fibonacci", log_n_rows=10, sequence_len=8), "smoke", 0, 1),
]


@pytest.fixture(scope="module")
def served(tmp_path_factory):
    d = tmp_path_factory.mktemp("session")
    frames = [prove_frame(i, w, str(d), proof_protocol=p, warmups=wu, samples=k)
              for i, (w, p, wu, k) in enumerate(REQUESTS)]
    frames.append(shutdown_frame(len(REQUESTS)))
    print("session: %d requests" % len(REQUESTS), flush=True)
    rc, out, err = run_session(frames, args=verifier_args(), timeout=110)
    assert rc == 0, (out[-1:], err[-2000:])
    reports = [json.load(open(f["outputs"]["report"])) for f in frames[:-1]]
    proofs = [open(f["outputs"]["proof"], "rb").read() for f in frames[:-1]]
    return dict(dir=d, frames=frames, out=out, reports=reports, proofs=proofs)


def _fixture_entry(workload):
    for e in load_fixture()["workloads"]:
        if e["workload"] == workload["example"] and all(workload[k] == v for k, v in e["parameters"].items()):
            return e
    return None


def test_session_frames(served):
    out = served["out"]
    assert [f["type"] for f in out] == ["ready"] + ["result"] * len(REQUESTS) + ["closed"]
    ready = out[0]
    assert ready["mode"] == "prove" and ready["device"]["arch"].startswith("gfx950")
    assert ready["verifier"]["required"] and ready["capabilities"]["verified_proofs"]
    for i, (f, p) in enumerate(zip(out[1:-1], served["proofs"])):
        assert f["sequence"] == i and f["request_id"] == "req-%d" % i
        assert f["status"] == "verified" and f["proof_verified"] and f["outputs_committed"]
        assert f["verified_proofs"] == REQUESTS[i][2] + REQUESTS[i][3]
        assert f["proof_bytes"] == len(p) and f["proof_sha256"] == hashlib.sha256(p).hexdigest()
    assert out[-1]["completed"] == len(REQUESTS)
    # the context outlives requests: later ones reuse the runtime, the arena and the tower
    assert not out[1]["reuse"]["runtime"] and all(f["reuse"]["runtime"] for f in out[2:-1])
    assert out[-2]["reuse"]["resident_arena"] and out[-2]["reuse"]["twiddle_tower"]
    # nothing but the published outputs is left in the output directory
    names = sorted(os.listdir(served["dir"]))
    assert names == sorted(["proof-%d.json" % i for i in range(len(REQUESTS))] +
                           ["report-%d.json" % i for i in range(len(REQUESTS))])


def test_session_proofs_are_the_reference_proofs(served):
    wf = {g["name"]: g for g in load_goldens()}["wide_fibonacci_log10_seq8_pow10"]
    po = {g["name"]: g for g in load_poseidon_goldens()}["poseidon_log10_pow10"]
    xo = {g["name"]: g for g in load_xor_goldens()}["xor_log14_step2_offset3_pow10"]
    p = served["proofs"]
    assert p[0] == wf["proof"] and p[1] == po["proof"] and p[2] == xo["proof"]
    for i in range(3):
        e = _fixture_entry(REQUESTS[i][0])
        r = served["reports"][i]
        assert hashlib.sha256(p[i]).hexdigest() == e["proof_sha256"] and len(p[i]) == e["proof_bytes"]
        assert r["workload"]["descriptor_sha256"] == e["descriptor_sha256"]
        for k in ("trace_log_rows", "committed_columns", "native_unit", "native_units"):
            assert r["workload"][k] == e[k], k
    assert p[3] == oracle.prove_logup_xor(16, 7, **FUNC)
    assert p[4] == oracle.prove_wide_fibonacci(log_n_rows=10, sequence_len=8, pow_bits=0, log_blowup_factor=1,
                                               log_last_layer_degree_bound=0, n_queries=3)


def _median_mad(v):
    m = statistics.median(v)
    return m, statistics.median([abs(x - m) for x in v])


def test_session_reports(served):
    for (w, proto, warm, k), r, p in zip(REQUESTS, served["reports"], served["proofs"]):
        assert r["schema_version"] == 7 and r["backend"] == "rocm_hip" and not r["profiled"]
        assert r["protocol"]["name"] == proto and r["protocol"]["fold_step"] == 1
        assert r["protocol"]["pow_bits"] == (0 if proto == "smoke" else 10)
        assert r["workload"]["name"] == w["example"]
        assert r["proof"]["verified_samples"] == k and r["proof"]["all_samples_byte_identical"]
        assert [s["sha256"] for s in r["proof"]["samples"]] == [hashlib.sha256(p).hexdigest()] * k
        assert r["proof"]["artifact"]["bytes"] == len(p)
        assert r["resources"]["measured_samples"] == k and r["resources"]["measured_warmups"] == warm
        assert len(r["timing"]["warmup_request_seconds"]) == warm
        # headline figures need 10 warmups and 3-5 samples: these requests are correctness runs
        assert r["evidence_class"] == "correctness_only" and not r["throughput"]["headline_eligible"]
        assert r["throughput"]["headline_native_mhz"] is None
        samples = r["timing"]["samples"]
        assert len(samples) == k
        for key in ("prove_seconds", "request_seconds", "verify_seconds", "proof_encode_seconds"):
            v = [s[key] for s in samples]
            summ = r["timing"][key]
            med, mad = _median_mad(v)
            assert summ["min"] == min(v) and summ["max"] == max(v)
            assert abs(summ["median"] - med) <= 1e-12 * max(1.0, med) and abs(summ["mad"] - mad) <= 1e-12 * max(1.0, mad)
        for s in samples:
            assert 0 < s["prove_seconds"] <= s["request_seconds"] and s["verify_seconds"] > 0
            assert abs(s["native_mhz"] - r["workload"]["native_units"] / s["prove_seconds"] / 1e6) <= 1e-9 * s["native_mhz"]
        # the estimate the request was admitted against bounds what it really held
        assert r["resource_admission"]["accounted_bytes"] >= r["resources"]["request_peak_device_bytes"] > 0
        assert r["resources"]["within_budget"]
        assert r["session"]["tower_build_count"] >= 1 and r["session"]["retained_device_twiddle_bytes"] > 0
        tele = r["backend_telemetry"]
        if w["example"] == "xor_logup":
            assert tele is None
        else:
            assert tele["valid"] and len(tele["samples"]) == k and tele["total_cpu_fallbacks"] == 0
            assert all(t["total_gpu_ms"] > 0 and len(t["stage_gpu_ms"]) == 8 for t in tele["samples"])


def test_session_report_layout_matches_the_reference(served):
    ref = load_fixture()["report_layout"]
    for r in served["reports"]:
        # product_identity is omitted when absent, as the reference's legacy encoding does
        assert set(r) == set(ref) - {"product_identity"}
        for sect in ("protocol", "workload", "timing", "throughput", "proof", "resources"):
            assert set(ref[sect]) <= set(r[sect]), (sect, set(ref[sect]) - set(r[sect]))


def test_session_without_verifier(tmp_path):
    frames = [prove_frame(0, dict(example="xor", log_size=12, log_step=1, offset=5), str(tmp_path)),
              shutdown_frame(1)]
    print("session without verifier", flush=True)
    rc, out, err = run_session(frames, timeout=110)
    assert rc == 0, err[-2000:]
    res = out[1]
    assert res["status"] == "unverified" and not res["proof_verified"] and res["verified_proofs"] == 0
    r = json.load(open(frames[0]["outputs"]["report"]))
    assert r["evidence_class"] == "unverified" and r["proof"]["verified_samples"] == 0
    proof = open(frames[0]["outputs"]["proof"], "rb").read()
    assert proof == oracle.prove_xor(12, 1, 5, **FUNC)


def test_session_fails_closed_on_a_bad_second_request(tmp_path):
    """A good request is published; the next, invalid one ends the session with an error frame
    and nothing of it is written (app.zig:194-205)."""
    good = prove_frame(0, dict(example="wide_fibonacci", log_n_rows=6, sequence_len=4), str(tmp_path))
    bad = prove_frame(1, dict(example="wide_fibonacci", log_n_rows=6, sequence_len=4), str(tmp_path),
                      budget="0.00001")
    print("session fail-closed", flush=True)
    rc, out, err = run_session([good, bad, shutdown_frame(2)], timeout=110)
    assert rc == 1
    assert [f["type"] for f in out] == ["ready", "result", "error"]
    assert out[-1]["code"] == "AccountedMemoryBudgetExceeded" and out[-1]["sequence"] == 1
    assert sorted(os.listdir(tmp_path)) == ["proof-0.json", "report-0.json"]
