"""The session front door on the CPU: its strict JSONL protocol (the checks of the reference's
src/tools/metal_session/protocol.zig:77-196 and the fail-closed loop of
src/tools/metal_prover_session/app.zig:150-211), run in --check mode, which validates and admits
every request without a device; the workload descriptor digests against the reference's own
schema-7 reports; the device-byte admission; and the external verifier command used by the GPU
session tests."""
import hashlib
import json
import os
import subprocess
import sys

import pytest

import stwo_rocm
from conftest import load_goldens
from session_util import (PROTOCOL, SESSION, VERIFY_CLI, descriptor, load_fixture, prove_frame, run_session,
                          shutdown_frame)

WF = {"example": "wide_fibonacci", "log_n_rows": 10, "sequence_len": 8}
FUNCTIONAL = dict(name="functional", pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3,
                  fold_step=1)


def _workload_of(e):
    return dict(example=e["workload"], **e["parameters"])


def _expect_error(frames, code, tmp_path=None, raw=None, args=("--check",)):
    rc, out, err = run_session(frames, args=args, raw_lines=raw)
    assert rc == 1, (out, err)
    assert out[0]["type"] == "ready"
    assert out[-1]["type"] == "error" and out[-1]["code"] == code, out[-1]
    assert out[-1]["protocol"] == PROTOCOL
    if tmp_path is not None:
        assert not [p for p in os.listdir(tmp_path) if p.startswith(("proof-", "report-"))]
    return out


def test_session_executable_built():
    assert os.access(SESSION, os.X_OK), "build the library first (python -c 'import __graft_entry__ as g; g.build()')"


def test_check_mode_accepts_a_sequence_and_closes(tmp_path):
    frames = [prove_frame(0, WF, str(tmp_path)),
              prove_frame(1, {"example": "poseidon", "log_n_instances": 10}, str(tmp_path)),
              prove_frame(2, {"example": "xor", "log_size": 14, "log_step": 2, "offset": 3}, str(tmp_path)),
              prove_frame(3, {"example": "xor_logup", "log_n": 16, "seed": 7}, str(tmp_path), proof_protocol="secure"),
              shutdown_frame(4)]
    rc, out, err = run_session(frames, args=["--check"])
    assert rc == 0, err
    assert [f["type"] for f in out] == ["ready", "checked", "checked", "checked", "checked", "closed"]
    ready = out[0]
    assert ready["mode"] == "check" and ready["device"] is None
    assert ready["capabilities"]["strict_order"] and ready["capabilities"]["atomic_outputs"]
    assert ready["library_sha256"] == hashlib.sha256(open(stwo_rocm.LIB_PATH, "rb").read()).hexdigest()
    assert ready["session_executable_sha256"] == hashlib.sha256(open(SESSION, "rb").read()).hexdigest()
    assert [f["sequence"] for f in out[1:5]] == [0, 1, 2, 3]
    assert out[-1]["completed"] == 4
    assert out[4]["descriptor"].endswith("|protocol=secure|pow_bits=26|log_blowup_factor=1|"
                                         "log_last_layer_degree_bound=0|n_queries=70|fold_step=1")
    # check mode writes nothing
    assert os.listdir(tmp_path) == []


def test_descriptor_digests_match_the_reference_reports(tmp_path):
    """Every wide-Fibonacci, Poseidon and XOR workload of the reference's schema-7 reports: the
    session's descriptor string hashes to the digest the reference reported for it."""
    fx = load_fixture()["workloads"]
    assert len(fx) >= 9
    frames = [prove_frame(i, _workload_of(e), str(tmp_path)) for i, e in enumerate(fx)] + [shutdown_frame(len(fx))]
    rc, out, err = run_session(frames, args=["--check"])
    assert rc == 0, err
    for e, f in zip(fx, out[1:-1]):
        assert e["protocol"] == FUNCTIONAL
        assert f["descriptor"] == descriptor(_workload_of(e), e["protocol"])
        assert f["descriptor_sha256"] == hashlib.sha256(f["descriptor"].encode()).hexdigest()
        assert f["descriptor_sha256"] == e["descriptor_sha256"], e["source"]


def test_admission_uses_the_serving_estimate(tmp_path):
    need = stwo_rocm.job_device_bytes(0, 10, 8)
    assert need == 16 * 1024 * 8 + 1024 * 1024 + (64 << 20)
    ok = prove_frame(0, WF, str(tmp_path), budget="%.9f" % (need / 2**30 + 1e-6))
    rc, out, err = run_session([ok, shutdown_frame(1)], args=["--check"])
    assert rc == 0, err
    assert out[1]["accounted_bytes"] == need and out[1]["budget_bytes"] >= need
    tight = prove_frame(0, WF, str(tmp_path), budget="%.9f" % ((need - 4096) / 2**30))
    _expect_error([tight, shutdown_frame(1)], "AccountedMemoryBudgetExceeded", tmp_path)
    big = prove_frame(0, {"example": "wide_fibonacci", "log_n_rows": 20, "sequence_len": 100}, str(tmp_path),
                      budget="1")
    _expect_error([big], "AccountedMemoryBudgetExceeded")


def _variants(tmp_path):
    d = str(tmp_path)
    base = prove_frame(0, WF, d)
    v = {}

    def mod(name, code, **kw):
        f = json.loads(json.dumps(base))
        for k, val in kw.items():
            if val is _DROP:
                del f[k]
            else:
                f[k] = val
        v[name] = (f, code)

    mod("wrong_protocol", "InvalidField", protocol="stwo-zig-metal-prover-session")
    mod("old_version", "InvalidField", version=0)
    mod("next_version", "InvalidField", version=2)
    mod("string_version", "InvalidField", version="1")
    mod("sequence_skips", "InvalidSequence", sequence=1)
    mod("negative_sequence", "InvalidSequence", sequence=-1)
    mod("extra_key", "InvalidFields", extra=1)
    mod("missing_key", "InvalidFields", samples=_DROP)
    mod("empty_request_id", "InvalidRequestId", request_id="")
    mod("request_id_symbol", "InvalidRequestId", request_id="-abc")
    mod("request_id_space", "InvalidRequestId", request_id="a b")
    mod("request_id_long", "InvalidRequestId", request_id="a" * 129)
    mod("budget_zero", "InvalidBudget", budget_gib="0")
    mod("budget_negative", "InvalidBudget", budget_gib="-2")
    mod("budget_text", "InvalidBudget", budget_gib="lots")
    mod("budget_inf", "InvalidBudget", budget_gib="inf")
    mod("budget_nan", "InvalidBudget", budget_gib="nan")
    mod("budget_number", "InvalidField", budget_gib=4)
    mod("unknown_example", "InvalidWorkload", workload={"example": "plonk", "log_n_rows": 10})
    mod("workload_extra", "InvalidWorkload", workload=dict(WF, seed=1))
    mod("log_rows_high", "InvalidWorkload", workload=dict(WF, log_n_rows=27))
    mod("seq_len_low", "InvalidWorkload", workload=dict(WF, sequence_len=2))
    mod("xor_step_over", "InvalidWorkload", workload={"example": "xor", "log_size": 10, "log_step": 11, "offset": 0})
    mod("logup_small", "InvalidWorkload", workload={"example": "xor_logup", "log_n": 15, "seed": 1})
    mod("float_log", "InvalidWorkload", workload=dict(WF, log_n_rows=10.0))
    mod("proof_protocol", "InvalidProofProtocol", proof_protocol="fast")
    mod("warmups_over", "InvalidField", warmups=31)
    mod("samples_zero", "InvalidField", samples=0)
    mod("samples_over", "InvalidField", samples=22)
    mod("relative_output", "InvalidPath", outputs={"proof": "proof.json", "report": d + "/r.json"})
    mod("same_outputs", "InvalidOutputs", outputs={"proof": d + "/x.json", "report": d + "/x.json"})
    mod("output_keys", "InvalidOutputs", outputs={"proof": d + "/x.json"})
    mod("missing_dir", "InvalidPath", outputs={"proof": d + "/nodir/p.json", "report": d + "/r.json"})
    mod("type_other", "InvalidFrame", type="status")
    return v


class _Drop:
    pass


_DROP = _Drop()


def test_protocol_rejections_fail_closed(tmp_path):
    """Each malformed request ends the session with one error frame naming the check, writes no
    output and leaves the exit status 1 (protocol.zig's strict parser; app.zig fails closed)."""
    for name, (frame, code) in _variants(tmp_path).items():
        out = _expect_error([frame, shutdown_frame(1)], code, tmp_path)
        assert out[-1]["sequence"] is None, name


def test_existing_output_is_refused(tmp_path):
    f = prove_frame(0, WF, str(tmp_path))
    open(f["outputs"]["report"], "w").write("keep")
    _expect_error([f, shutdown_frame(1)], "OutputAlreadyExists")
    assert open(f["outputs"]["report"]).read() == "keep"


def test_frame_level_failures(tmp_path):
    good = json.dumps(prove_frame(0, WF, str(tmp_path)), separators=(",", ":"))
    shut = json.dumps(shutdown_frame(1))
    cases = {
        "UncleanEndOfStream": [good],
        "EmptyFrame": [good, ""],
        "InvalidFrame": ["{not json"],
    }
    for code, raw in cases.items():
        _expect_error(None, code, raw=raw)
    # JSON strictness: trailing text, duplicate keys, leading zeros, lone surrogates, deep nesting
    for bad in [good + " x", good[:-1] + ',"type":"prove"}', good.replace('"sequence":0', '"sequence":00'),
                good.replace('"req-0"', '"\\ud800"'), "[" * 70 + "]" * 70, "[]", '"prove"']:
        _expect_error(None, "InvalidFrame", raw=[bad, shut])
    # shutdown must name the next sequence, carry exactly its four fields, and end the session
    rc, out, _ = run_session([shutdown_frame(0)], args=["--check"])
    assert rc == 0 and out[-1] == {"protocol": PROTOCOL, "version": 1, "type": "closed", "completed": 0}
    _expect_error([prove_frame(0, WF, str(tmp_path)), shutdown_frame(0)], "InvalidSequence")
    _expect_error([dict(shutdown_frame(0), extra=1)], "InvalidShutdown")
    rc, out, _ = run_session([shutdown_frame(0), "garbage after shutdown"], args=["--check"])
    assert rc == 0 and out[-1]["type"] == "closed"
    # a frame over 4 MiB is refused before it is parsed
    _expect_error(None, "FrameTooLarge", raw=["{" + " " * (4 << 20) + "}"])


def test_request_ids_and_unicode_accepted(tmp_path):
    frames = [prove_frame(0, WF, str(tmp_path), request_id="A.b_c:d-9"),
              prove_frame(1, WF, str(tmp_path), request_id="x" * 128), shutdown_frame(2)]
    # escapes and non-ASCII text in a string field that is not otherwise constrained
    raw = [json.dumps(f) for f in frames]
    raw[0] = raw[0].replace('"functional"', '"\\u0066unctional"')
    rc, out, err = run_session(None, args=["--check"], raw_lines=raw)
    assert rc == 0, err
    assert [f["request_id"] for f in out[1:3]] == ["A.b_c:d-9", "x" * 128]


def test_bad_command_lines():
    for args in (["--verifier-arg", "x"], ["--bogus"]):
        r = subprocess.run([SESSION] + args, input="", capture_output=True, text=True, timeout=60)
        assert r.returncode == 2
    rc, out, _ = run_session([shutdown_frame(0)], args=["--check", "--verifier", "relative/verifier"])
    assert rc == 1 and out[-1]["code"] == "VerifierUnavailable"


def test_verifier_command_on_golden_proofs(tmp_path):
    """The external verifier command the GPU session tests hand to the session: accepts a
    reference golden proof and rejects it under a different statement."""
    e = {g["name"]: g for g in load_goldens()}["wide_fibonacci_log10_seq8_pow10"]
    proof = tmp_path / "p.json"
    proof.write_bytes(e["proof"])
    pc = e["pcs_config"]
    cfg = dict(pow_bits=pc["pow_bits"], log_blowup_factor=pc["fri_config"]["log_blowup_factor"],
               log_last_layer_degree_bound=pc["fri_config"]["log_last_layer_degree_bound"],
               n_queries=pc["fri_config"]["n_queries"])
    good = tmp_path / "s.json"
    good.write_text(json.dumps(dict(air="wide_fibonacci", statement=e["statement"], pcs_config=cfg)))
    bad = tmp_path / "b.json"
    bad.write_text(json.dumps(dict(air="wide_fibonacci", statement=dict(e["statement"], sequence_len=9),
                                   pcs_config=cfg)))
    ok = subprocess.run([sys.executable, VERIFY_CLI, "--statement", str(good), "--proof", str(proof)],
                        capture_output=True, text=True, timeout=300)
    assert ok.returncode == 0, ok.stderr
    no = subprocess.run([sys.executable, VERIFY_CLI, "--statement", str(bad), "--proof", str(proof)],
                        capture_output=True, text=True, timeout=300)
    assert no.returncode == 1 and "rejected" in no.stderr
