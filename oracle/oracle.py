"""ctypes binding of the C oracle -- TEST INFRASTRUCTURE ONLY.

The oracle (oracle/stwo_oracle.c) is a CPU restatement of the reference's CPU
backend used to check the HIP backend.  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libstwo_oracle.so")
P = (1 << 31) - 1

_lib = None


class WideFibParams(ctypes.Structure):
    _fields_ = [
        ("log_n_rows", ctypes.c_uint32),
        ("sequence_len", ctypes.c_uint32),
        ("pow_bits", ctypes.c_uint32),
        ("log_blowup_factor", ctypes.c_uint32),
        ("log_last_layer_degree_bound", ctypes.c_uint32),
        ("n_queries", ctypes.c_uint32),
        ("trace_mode", ctypes.c_uint32),
        ("trace_seed", ctypes.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.ora_m31_mul.restype = ctypes.c_uint32
        L.ora_m31_mul.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.ora_m31_inv.restype = ctypes.c_uint32
        L.ora_m31_inv.argtypes = [ctypes.c_uint32]
        L.ora_qm31_mul.argtypes = [u32p, u32p, u32p]
        L.ora_qm31_inv.argtypes = [u32p, u32p]
        L.ora_circle_point_at_index.argtypes = [ctypes.c_uint32, u32p]
        L.ora_circle_domain_point.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.ora_twiddles.argtypes = [ctypes.c_uint32, u32p, u32p]
        L.ora_circle_evaluate.argtypes = [u32p, ctypes.c_uint32]
        L.ora_circle_interpolate.argtypes = [u32p, ctypes.c_uint32]
        L.ora_circle_lde.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p]
        L.ora_eval_at_point.argtypes = [u32p, ctypes.c_uint32, u32p, u32p]
        L.ora_blake2s.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_char_p]
        L.ora_merkle_commit.restype = ctypes.c_void_p
        L.ora_merkle_commit.argtypes = [ctypes.POINTER(u32p), u32p, ctypes.c_uint32]
        L.ora_tree_root.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.ora_tree_max_log.restype = ctypes.c_uint32
        L.ora_tree_max_log.argtypes = [ctypes.c_void_p]
        L.ora_tree_layer.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_char_p]
        L.ora_tree_free.argtypes = [ctypes.c_void_p]
        L.ora_fold_line.argtypes = [u32p, ctypes.c_uint32, ctypes.c_uint32, u32p, u32p]
        L.ora_fold_circle_into_line.argtypes = [ctypes.POINTER(u32p), ctypes.c_uint32, u32p, u32p]
        L.ora_pow_grind.restype = ctypes.c_uint64
        L.ora_pow_grind.argtypes = [ctypes.c_char_p, ctypes.c_uint32]
        L.ora_widefib_trace.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p]
        L.ora_prove_wide_fibonacci.restype = ctypes.c_int
        L.ora_prove_wide_fibonacci.argtypes = [ctypes.POINTER(WideFibParams), ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t)]
        L.ora_last_timings.argtypes = [ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def m31_mul(a, b):
    return lib().ora_m31_mul(a, b)


def m31_inv(a):
    return lib().ora_m31_inv(a)


def qm31_mul(a, b):
    a, pa = _u32(a)
    b, pb = _u32(b)
    o, po = _u32(np.zeros(4))
    lib().ora_qm31_mul(pa, pb, po)
    return [int(x) for x in o]


def qm31_inv(a):
    a, pa = _u32(a)
    o, po = _u32(np.zeros(4))
    lib().ora_qm31_inv(pa, po)
    return [int(x) for x in o]


def circle_point_at_index(idx):
    o, po = _u32(np.zeros(2))
    lib().ora_circle_point_at_index(idx, po)
    return int(o[0]), int(o[1])


def circle_domain_point(log_size, i):
    o, po = _u32(np.zeros(2))
    lib().ora_circle_domain_point(log_size, i, po)
    return int(o[0]), int(o[1])


def twiddles(log):
    n = 1 << (log - 1)
    t, pt = _u32(np.zeros(n))
    it, pit = _u32(np.zeros(n))
    lib().ora_twiddles(log, pt, pit)
    return t, it


def circle_evaluate(coeffs, log):
    v, pv = _u32(np.array(coeffs, dtype=np.uint32).copy())
    lib().ora_circle_evaluate(pv, log)
    return v


def circle_interpolate(values, log):
    v, pv = _u32(np.array(values, dtype=np.uint32).copy())
    lib().ora_circle_interpolate(pv, log)
    return v


def circle_lde(coeffs, log, log_blowup):
    c, pc = _u32(coeffs)
    o, po = _u32(np.zeros(1 << (log + log_blowup)))
    lib().ora_circle_lde(pc, log, log_blowup, po)
    return o


def eval_at_point(coeffs, log, point8):
    c, pc = _u32(coeffs)
    p, pp = _u32(point8)
    o, po = _u32(np.zeros(4))
    lib().ora_eval_at_point(pc, log, pp, po)
    return [int(x) for x in o]


def blake2s(data: bytes) -> bytes:
    out = ctypes.create_string_buffer(32)
    lib().ora_blake2s(data, len(data), out)
    return out.raw


def merkle_commit(columns):
    """columns: list of uint32 arrays (power-of-two lengths >= 2). Returns (root, layers)."""
    arrs = [np.ascontiguousarray(c, dtype=np.uint32) for c in columns]
    n = len(arrs)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    ptrs = (u32p * max(n, 1))(*[a.ctypes.data_as(u32p) for a in arrs])
    logs = np.array([int(a.size).bit_length() - 1 for a in arrs] or [0], dtype=np.uint32)
    _, plogs = _u32(logs)
    t = lib().ora_merkle_commit(ptrs, plogs, n)
    root = ctypes.create_string_buffer(32)
    lib().ora_tree_root(t, root)
    max_log = lib().ora_tree_max_log(t)
    layers = []
    for k in range(max_log + 1):
        buf = ctypes.create_string_buffer(32 << k)
        lib().ora_tree_layer(t, k, buf)
        layers.append(buf.raw)
    lib().ora_tree_free(t)
    return root.raw, layers


def fold_line(src_qm31, log_src, coset_initial_index, alpha):
    s, ps = _u32(src_qm31)
    a, pa = _u32(alpha)
    o, po = _u32(np.zeros(2 << log_src))
    lib().ora_fold_line(ps, log_src, coset_initial_index, pa, po)
    return o


def fold_circle_into_line(coords, log_src, alpha, dst=None):
    u32p = ctypes.POINTER(ctypes.c_uint32)
    arrs = [np.ascontiguousarray(c, dtype=np.uint32) for c in coords]
    ptrs = (u32p * 4)(*[a.ctypes.data_as(u32p) for a in arrs])
    a, pa = _u32(alpha)
    o, po = _u32(np.zeros(2 << log_src) if dst is None else np.array(dst, dtype=np.uint32).copy())
    lib().ora_fold_circle_into_line(ptrs, log_src, pa, po)
    return o


def pow_grind(digest: bytes, bits: int) -> int:
    return lib().ora_pow_grind(digest, bits)


def widefib_trace(log_n, seq_len, mode=0, seed=0):
    o, po = _u32(np.zeros(seq_len << log_n))
    lib().ora_widefib_trace(log_n, seq_len, mode, seed, po)
    return o.reshape(seq_len, 1 << log_n)


def prove_wide_fibonacci(log_n_rows, sequence_len, pow_bits=10, log_blowup_factor=1,
                         log_last_layer_degree_bound=0, n_queries=3, trace_mode=0, trace_seed=0):
    prm = WideFibParams(log_n_rows, sequence_len, pow_bits, log_blowup_factor,
                        log_last_layer_degree_bound, n_queries, trace_mode, trace_seed)
    cap = 1 << 20
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t(0)
        rc = lib().ora_prove_wide_fibonacci(ctypes.byref(prm), buf, cap, ctypes.byref(n))
        if rc == 2:
            cap = n.value + 1
            continue
        if rc != 0:
            raise RuntimeError("oracle prove failed with code %d" % rc)
        return buf.raw[: n.value]


def last_timings():
    arr = (ctypes.c_double * 8)()
    lib().ora_last_timings(arr)
    keys = ["trace_gen", "main_commit", "composition", "sampled_values", "quotients", "fri_commit", "pow", "prove_total"]
    return dict(zip(keys, list(arr)))
