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
        L.ora_poseidon_trace.argtypes = [ctypes.c_uint32, u32p]
        L.ora_prove_xor.restype = ctypes.c_int
        L.ora_prove_xor.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(WideFibParams),
                                    ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_xor_example_columns.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p]
        L.ora_prove_logup_xor.restype = ctypes.c_int
        L.ora_prove_logup_xor.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(WideFibParams),
                                          ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_prove_poseidon.restype = ctypes.c_int
        L.ora_prove_poseidon.argtypes = [ctypes.c_uint32, ctypes.POINTER(WideFibParams), ctypes.c_char_p,
                                         ctypes.c_size_t, ctypes.POINTER(ctypes.c_size_t)]
        L.ora_prove_wide_fibonacci.restype = ctypes.c_int
        L.ora_prove_wide_fibonacci.argtypes = [ctypes.POINTER(WideFibParams), ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t)]
        L.ora_last_timings.argtypes = [ctypes.POINTER(ctypes.c_double)]
        L.ora_set_threads.argtypes = [ctypes.c_int]
        L.ora_get_threads.restype = ctypes.c_int
        L.ora_xor_trace.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, u32p, u32p, u32p]
        L.ora_logup_denominators.argtypes = [u32p, u32p, u32p, ctypes.c_uint64, u32p, u32p, u32p]
        L.ora_logup_cumulative.restype = ctypes.c_int
        L.ora_logup_cumulative.argtypes = [ctypes.c_int, u32p, u32p, ctypes.c_uint64, u32p, u32p]
        L.ora_gkr_gen_eq_evals.argtypes = [u32p, ctypes.c_uint32, u32p, u32p]
        L.ora_gkr_next_layer.argtypes = [ctypes.c_int, u32p, u32p, ctypes.c_uint32, u32p, u32p]
        L.ora_mle_fix_first.argtypes = [ctypes.c_int, u32p, ctypes.c_uint32, u32p, u32p]
        L.ora_gkr_prove.restype = ctypes.c_size_t
        L.ora_gkr_prove.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(u32p),
                                    ctypes.POINTER(u32p), u32p, ctypes.c_char_p, u32p, ctypes.c_size_t]
        L.ora_batched_pipeline.restype = ctypes.c_size_t
        L.ora_batched_pipeline.argtypes = [ctypes.c_uint32] * 6 + [ctypes.c_uint64, u32p, ctypes.c_size_t]
        L.ora_logup_xor_pipeline.restype = ctypes.c_size_t
        L.ora_logup_xor_pipeline.argtypes = [ctypes.c_uint32, ctypes.c_uint64, u32p, ctypes.c_size_t]
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


def poseidon_trace(log_n_rows):
    o, po = _u32(np.zeros(1264 << log_n_rows))
    lib().ora_poseidon_trace(log_n_rows, po)
    return o.reshape(1264, 1 << log_n_rows)


def prove_poseidon(log_n_instances, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3):
    prm = WideFibParams(0, 0, pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries, 0, 0)
    cap = 1 << 20
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t(0)
        rc = lib().ora_prove_poseidon(log_n_instances, ctypes.byref(prm), buf, cap, ctypes.byref(n))
        if rc == 2:
            cap = n.value + 1
            continue
        if rc != 0:
            raise RuntimeError("oracle prove failed with code %d" % rc)
        return buf.raw[: n.value]


def xor_example_columns(log_size, log_step, offset):
    """(is_first, is_step_with_offset, main) of the periodic XOR example AIR."""
    cols = [np.zeros(1 << log_size, dtype=np.uint32) for _ in range(3)]
    ptrs = [_u32(c)[1] for c in cols]
    lib().ora_xor_example_columns(log_size, log_step, offset, *ptrs)
    return cols


def prove_xor(log_size, log_step, offset, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0,
              n_queries=3):
    prm = WideFibParams(0, 0, pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries, 0, 0)
    cap = 1 << 20
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t(0)
        rc = lib().ora_prove_xor(log_size, log_step, offset, ctypes.byref(prm), buf, cap, ctypes.byref(n))
        if rc == 2:
            cap = n.value + 1
            continue
        if rc != 0:
            raise RuntimeError("oracle prove failed with code %d" % rc)
        return buf.raw[: n.value]


def prove_logup_xor(log_n, seed, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3):
    """Whole proof of the XOR-lookup AIR (see ora_prove_logup_xor)."""
    prm = WideFibParams(0, 0, pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries, 0, 0)
    cap = 1 << 20
    while True:
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t(0)
        rc = lib().ora_prove_logup_xor(log_n, seed, ctypes.byref(prm), buf, cap, ctypes.byref(n))
        if rc == 2:
            cap = n.value + 1
            continue
        if rc != 0:
            raise RuntimeError("oracle prove failed with code %d" % rc)
        return buf.raw[: n.value]


def set_threads(n):
    """OpenMP thread count of the following oracle calls."""
    lib().ora_set_threads(int(n))


def get_threads():
    return int(lib().ora_get_threads())


def last_timings():
    arr = (ctypes.c_double * 8)()
    lib().ora_last_timings(arr)
    keys = ["trace_gen", "main_commit", "composition", "sampled_values", "quotients", "fri_commit", "pow", "prove_total"]
    return dict(zip(keys, list(arr)))


# ---------------------------------------------------------------- LogUp / GKR
GKR_GRAND_PRODUCT, GKR_LOGUP_GENERIC, GKR_LOGUP_MULTIPLICITIES, GKR_LOGUP_SINGLES = 0, 1, 2, 3
_NULL = ctypes.POINTER(ctypes.c_uint32)()


def xor_trace(log_size, seed):
    """(a, b, c, multiplicities[65536]) of the XOR-lookup AIR."""
    n = 1 << log_size
    a, b, c = (np.zeros(n, np.uint32) for _ in range(3))
    m = np.zeros(1 << 16, np.uint32)
    _, pa = _u32(a)
    _, pb = _u32(b)
    _, pc = _u32(c)
    _, pm = _u32(m)
    lib().ora_xor_trace(log_size, seed, pa, pb, pc, pm)
    return a, b, c, m


def logup_denominators(a, b, c, z, alpha, n=None):
    """QM31 AoS denominators z - (a + b*alpha + c*alpha^2); table rows (x, y, x^y) of index x*256+y when a is None."""
    zz, pz = _u32(z)
    al, pal = _u32(alpha)
    if a is None:
        n = 1 << 16 if n is None else n
        out = np.zeros(4 * n, np.uint32)
        _, po = _u32(out)
        lib().ora_logup_denominators(_NULL, _NULL, _NULL, n, pz, pal, po)
        return out
    a, pa = _u32(a)
    b, pb = _u32(b)
    c, pc = _u32(c)
    out = np.zeros(4 * len(a), np.uint32)
    _, po = _u32(out)
    lib().ora_logup_denominators(pa, pb, pc, len(a), pz, pal, po)
    return out


def logup_cumulative(num_kind, num, den):
    """(sums QM31 AoS, claimed[4]); num_kind 0 = QM31, 1 = M31, 2 = ones (num ignored)."""
    d, pd = _u32(den)
    n = len(d) // 4
    sums = np.zeros(4 * n, np.uint32)
    claimed = np.zeros(4, np.uint32)
    _, ps = _u32(sums)
    _, pcl = _u32(claimed)
    if num is None:
        pn = _NULL
    else:
        num, pn = _u32(num)
    rc = lib().ora_logup_cumulative(num_kind, pn, pd, n, ps, pcl)
    if rc != 0:
        raise ZeroDivisionError("zero LogUp denominator")
    return sums, claimed


def gkr_gen_eq_evals(y, scale):
    """y: QM31 AoS (n points); returns 2^n QM31 AoS."""
    y, py = _u32(y)
    n = len(y) // 4
    sc, psc = _u32(scale)
    out = np.zeros(4 << n, np.uint32)
    _, po = _u32(out)
    lib().ora_gkr_gen_eq_evals(py, n, psc, po)
    return out


def gkr_next_layer(kind, num, den, n_vars):
    d, pd = _u32(den)
    pn = _NULL
    if num is not None:
        num, pn = _u32(num)
    half = 1 << (n_vars - 1)
    on = np.zeros(4 * half, np.uint32)
    od = np.zeros(4 * half, np.uint32)
    _, pon = _u32(on)
    _, pod = _u32(od)
    lib().ora_gkr_next_layer(kind, pn, pd, n_vars, pon, pod)
    return (None if kind == GKR_GRAND_PRODUCT else on), od


def mle_fix_first(base, values, n_vars, assignment):
    v, pv = _u32(values)
    a, pa = _u32(assignment)
    out = np.zeros(4 << (n_vars - 1), np.uint32)
    _, po = _u32(out)
    lib().ora_mle_fix_first(1 if base else 0, pv, n_vars, pa, po)
    return out


def gkr_prove(instances, digest=bytes(32)):
    """instances: list of (kind, num or None, den, n_vars).  Returns the proof word stream."""
    n = len(instances)
    kinds = (ctypes.c_int * n)(*[k for k, _, _, _ in instances])
    keep = []
    nums = (ctypes.POINTER(ctypes.c_uint32) * n)()
    dens = (ctypes.POINTER(ctypes.c_uint32) * n)()
    for i, (k, num, den, nv) in enumerate(instances):
        d, pd = _u32(den)
        keep.append(d)
        dens[i] = pd
        if num is not None:
            nn, pn = _u32(num)
            keep.append(nn)
            nums[i] = pn
        else:
            nums[i] = _NULL
    nv, pnv = _u32([nv for _, _, _, nv in instances])
    need = lib().ora_gkr_prove(n, kinds, nums, dens, pnv, digest, _NULL, 0)
    out = np.zeros(need, np.uint32)
    _, po = _u32(out)
    lib().ora_gkr_prove(n, kinds, nums, dens, pnv, digest, po, need)
    return out


def logup_xor_pipeline(log_n, seed):
    """Word stream of the LogUp workload (see ora_logup_xor_pipeline)."""
    need = lib().ora_logup_xor_pipeline(log_n, seed, _NULL, 0)
    if need == 0:
        raise ValueError("LogUp workload did not balance")
    out = np.zeros(need, np.uint32)
    _, po = _u32(out)
    lib().ora_logup_xor_pipeline(log_n, seed, po, need)
    return out


def batched_pipeline(n_columns, log_min, log_max, log_blowup, fri_log, n_folds, seed):
    """Word stream of the batched primitive pipeline (see ora_batched_pipeline)."""
    args = (n_columns, log_min, log_max, log_blowup, fri_log, n_folds, seed)
    need = lib().ora_batched_pipeline(*args, _NULL, 0)
    out = np.zeros(need, np.uint32)
    _, po = _u32(out)
    lib().ora_batched_pipeline(*args, po, need)
    return out
