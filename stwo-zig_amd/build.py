"""Build librocm backend: hipcc for gfx950, in-tree output stwo-zig_amd/lib/libstwo_rocm.so.

Each HIP source is compiled to an object in parallel, then linked into one
shared library that exports the C-ABI of include/stwo_rocm.h.  The session
front door (stwo-zig_amd/lib/stwo_rocm_session) is host C++ linked against it.
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
SESSION_SRC = os.path.join(HERE, "session", "stwo_rocm_session.cpp")
SESSION = os.path.join(LIB_DIR, "stwo_rocm_session")
SOURCES = ["runtime.hip", "kernels_fft.hip", "kernels_merkle.hip", "kernels_fri.hip", "kernels_prove.hip",
           "kernels_misc.hip", "kernels_chan.hip", "kernels_logup.hip", "gkr.hip", "transcript.hip", "logup_pipeline.hip",
           "batched_pipeline.hip", "serving.hip",
           "prover.hip", "hooks.hip", "logup_proof.hip", "batch_prover.hip",
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
    build_session()
    return LIB


def build_session():
    """The JSONL session front door: host C++ over the C-ABI, linked against the library beside it."""
    tmp = SESSION + ".tmp"
    root = os.path.dirname(HERE)
    commit, dirty = "unknown", 1
    r = subprocess.run(["git", "-C", root, "rev-parse", "HEAD"], capture_output=True, text=True)
    if r.returncode == 0:
        commit = r.stdout.strip()
        st = subprocess.run(["git", "-C", root, "status", "--porcelain", "--untracked-files=no"], capture_output=True,
                            text=True)
        dirty = int(st.returncode != 0 or bool(st.stdout.strip()))
    cmd = [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-DNDEBUG", "-Wall", "-Wextra", "-Wno-unused-parameter",
           '-DSTWO_ROCM_GIT_COMMIT="%s"' % commit, "-DSTWO_ROCM_GIT_DIRTY=%d" % dirty,
           SESSION_SRC, "-o", tmp, "-L" + LIB_DIR, "-lstwo_rocm", "-Wl,-rpath,$ORIGIN", "-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("session build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    if r.stderr.strip():
        print(r.stderr, file=sys.stderr)
    os.replace(tmp, SESSION)
    return SESSION


if __name__ == "__main__":
    print(build())
