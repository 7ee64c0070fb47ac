import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "stwo-zig_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN_DIR = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs the rocm C-ABI")
    config.addinivalue_line("markers", "slow: long-running CPU case")


@pytest.fixture(scope="session")
def ctx():
    """One HIP context for the whole GPU session (one process on the card)."""
    import stwo_rocm

    c = stwo_rocm.Context(0)
    yield c
    c.close()


def load_goldens():
    import json

    with open(os.path.join(GOLDEN_DIR, "wide_fibonacci_goldens.json")) as f:
        idx = json.load(f)
    out = []
    for name, e in sorted(idx.items(), key=lambda kv: (kv[1]["statement"]["log_n_rows"], kv[0])):
        with open(os.path.join(GOLDEN_DIR, e["file"]), "rb") as f:
            e = dict(e, name=name, proof=f.read())
        out.append(e)
    return out


def _example_goldens(example):
    import json

    with open(os.path.join(GOLDEN_DIR, example + "_goldens.json")) as f:
        idx = json.load(f)
    out = []
    for name, e in sorted(idx.items()):
        with open(os.path.join(GOLDEN_DIR, e["file"]), "rb") as f:
            out.append(dict(e, name=name, proof=f.read()))
    return out


def load_poseidon_goldens():
    """Poseidon AIR proofs of the reference CPU backend (benchmark rows log_n_instances 10, 13)."""
    return _example_goldens("poseidon")


def load_xor_goldens():
    """Periodic XOR example AIR proofs of the reference CPU backend (log_size 14, 16; step 4, offset 3)."""
    return _example_goldens("xor")


def xor_kwargs(e):
    c, st = e["pcs_config"], e["statement"]
    return dict(log_size=st["log_size"], log_step=st["log_step"], offset=st["offset"], pow_bits=c["pow_bits"],
                log_blowup_factor=c["fri_config"]["log_blowup_factor"],
                log_last_layer_degree_bound=c["fri_config"]["log_last_layer_degree_bound"],
                n_queries=c["fri_config"]["n_queries"])


def poseidon_kwargs(e):
    c = e["pcs_config"]
    return dict(log_n_instances=e["statement"]["log_n_instances"], pow_bits=c["pow_bits"],
                log_blowup_factor=c["fri_config"]["log_blowup_factor"],
                log_last_layer_degree_bound=c["fri_config"]["log_last_layer_degree_bound"],
                n_queries=c["fri_config"]["n_queries"])


def golden_kwargs(e):
    c = e["pcs_config"]
    st = e["statement"]
    return dict(log_n_rows=st["log_n_rows"], sequence_len=st["sequence_len"], pow_bits=c["pow_bits"],
                log_blowup_factor=c["fri_config"]["log_blowup_factor"],
                log_last_layer_degree_bound=c["fri_config"]["log_last_layer_degree_bound"],
                n_queries=c["fri_config"]["n_queries"])
