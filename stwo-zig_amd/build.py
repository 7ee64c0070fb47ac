"""Build librocm backend: hipcc for gfx950, in-tree output stwo-zig_amd/lib/libstwo_rocm.so.

Each HIP source is compiled to an object in parallel, then linked into one
shared library that exports the C-ABI of include/stwo_rocm.h.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libstwo_rocm.so")
SOURCES = ["runtime.hip", "kernels_fft.hip", "kernels_merkle.hip", "kernels_fri.hip", "kernels_prove.hip",
           "kernels_misc.hip", "kernels_chan.hip", "kernels_logup.hip", "gkr.hip", "transcript.hip", "logup_pipeline.hip",
           "batched_pipeline.hip", "serving.hip",
           "prover.hip", "hooks.hip", "logup_proof.hip",
           "capi.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("STWO_ROCM_ARCH", "gfx950")
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Wno-unused-variable", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(src):
    out = os.path.join(OBJ, src.rsplit(".", 1)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    # keep real diagnostics, drop the host-pass note about the packed-fp32 feature
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return out, warn


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, w in results:
        if w and verbose:
            print(w, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + [o for o, _ in results]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())
