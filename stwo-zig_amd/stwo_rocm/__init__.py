"""Python binding of the rocm proving backend (ctypes over include/stwo_rocm.h).

The method names follow the reference backend hooks a Zig ``RocmBackend`` would
forward to the same C-ABI (src/backend/*.zig contracts, the MetalCommitBackend
hook set in src/backends/metal/commit_backend.zig): interpolate/evaluate circle
buffers, commit_merkle, fold_line, fold_circle_into_line, compute quotients,
grind, and the whole wide-Fibonacci proof.

There is no fallback: constructing a ``Context`` without the built library or
without a HIP device raises ``RocmError``.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libstwo_rocm.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "stwo_rocm.h")
P = (1 << 31) - 1

_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p
_lib = None


class RocmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("stwo_rocm error %d: %s" % (code, msg))
        self.code = code


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


class PoseidonParams(ctypes.Structure):
    _fields_ = [
        ("log_n_instances", ctypes.c_uint32),
        ("pow_bits", ctypes.c_uint32),
        ("log_blowup_factor", ctypes.c_uint32),
        ("log_last_layer_degree_bound", ctypes.c_uint32),
        ("n_queries", ctypes.c_uint32),
    ]


class XorParams(ctypes.Structure):
    _fields_ = [
        ("log_size", ctypes.c_uint32),
        ("log_step", ctypes.c_uint32),
        ("offset", ctypes.c_uint64),
        ("pow_bits", ctypes.c_uint32),
        ("log_blowup_factor", ctypes.c_uint32),
        ("log_last_layer_degree_bound", ctypes.c_uint32),
        ("n_queries", ctypes.c_uint32),
    ]


class LogupParams(ctypes.Structure):
    _fields_ = [
        ("log_n", ctypes.c_uint32),
        ("pow_bits", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("log_blowup_factor", ctypes.c_uint32),
        ("log_last_layer_degree_bound", ctypes.c_uint32),
        ("n_queries", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class ProveStats(ctypes.Structure):
    _fields_ = [
        ("stage_gpu_ms", ctypes.c_double * 8),
        ("stage_host_ms", ctypes.c_double * 8),
        ("total_gpu_ms", ctypes.c_double),
        ("total_host_ms", ctypes.c_double),
        ("peak_hbm_bytes", ctypes.c_uint64),
    ]

    STAGES = ["main_commit", "composition", "sampled_values", "quotients", "fri_commit", "pow", "decommit", "encode"]

    def as_dict(self):
        return {
            "stage_gpu_ms": dict(zip(self.STAGES, list(self.stage_gpu_ms))),
            "stage_host_ms": dict(zip(self.STAGES, list(self.stage_host_ms))),
            "total_gpu_ms": self.total_gpu_ms,
            "total_host_ms": self.total_host_ms,
            "peak_hbm_bytes": self.peak_hbm_bytes,
        }


# (name, restype, argtypes) for every entry point of include/stwo_rocm.h
_SIGNATURES = [
    ("stwo_rocm_ctx_create", ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(_vp), ctypes.c_char_p,
                                             ctypes.c_size_t]),
    ("stwo_rocm_ctx_destroy", ctypes.c_int, [_vp]),
    ("stwo_rocm_last_error", ctypes.c_char_p, [_vp]),
    ("stwo_rocm_ctx_identity", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t]),
    ("stwo_rocm_synchronize", ctypes.c_int, [_vp]),
    ("stwo_rocm_memory_info", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                              ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_release_cached", ctypes.c_int, [_vp]),
    ("stwo_rocm_set_budget", ctypes.c_int, [_vp, ctypes.c_uint64]),
    ("stwo_rocm_get_runtime_info", ctypes.c_int, [_vp, ctypes.c_int, _vp]),
    ("stwo_rocm_job_device_bytes", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_buffer_alloc", ctypes.c_int, [_vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("stwo_rocm_buffer_free", ctypes.c_int, [_vp, _vp]),
    ("stwo_rocm_buffer_upload", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64]),
    ("stwo_rocm_buffer_download", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64]),
    ("stwo_rocm_twiddles_prepare", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("stwo_rocm_twiddles_read", ctypes.c_int, [_vp, ctypes.c_uint32, _u32p, _u32p]),
    ("stwo_rocm_batch_inverse", ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp]),
    ("stwo_rocm_bit_reverse", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32]),
    ("stwo_rocm_accumulate", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64]),
    ("stwo_rocm_circle_interpolate", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32,
                                                     ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_circle_evaluate", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, ctypes.POINTER(_vp),
                                                  ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_eval_at_point", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32, _u32p,
                                                _u32p]),
    ("stwo_rocm_merkle_commit", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_uint32, ctypes.POINTER(_vp),
                                                ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_merkle_parent_layer", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp]),
    ("stwo_rocm_tree_root", ctypes.c_int, [_vp, _vp, ctypes.c_char_p]),
    ("stwo_rocm_tree_max_log", ctypes.c_int, [_vp, _u32p]),
    ("stwo_rocm_tree_read_hashes", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, ctypes.c_uint32,
                                                   ctypes.c_char_p]),
    ("stwo_rocm_tree_decommit", ctypes.c_int, [_vp, _vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                                ctypes.c_char_p, ctypes.c_uint32, _u32p]),
    ("stwo_rocm_tree_free", ctypes.c_int, [_vp, _vp]),
    ("stwo_rocm_fri_fold_line", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _vp]),
    ("stwo_rocm_fri_fold_circle_into_line", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, _u32p, _vp]),
    ("stwo_rocm_compute_quotients", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(_vp), _u32p, ctypes.c_uint32,
                                                    _u32p, _u32p, _u32p, ctypes.c_uint32, _u32p,
                                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    ("stwo_rocm_widefib_trace", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint64, _vp]),
    ("stwo_rocm_widefib_composition", ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.c_uint32, ctypes.c_uint32,
                                                      _u32p, ctypes.POINTER(_vp)]),
    ("stwo_rocm_pow_grind", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_prove_xor", ctypes.c_int, [_vp, ctypes.POINTER(XorParams), ctypes.c_char_p, ctypes.c_uint64,
                                            ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ProveStats)]),
    ("stwo_rocm_prove_wide_fibonacci_batch", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _vp, _vp, ctypes.c_uint64,
                                                             _vp, _vp]),
    ("stwo_rocm_batch_prover_create", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("stwo_rocm_batch_prover_run", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp, ctypes.c_uint64, _vp, _vp]),
    ("stwo_rocm_batch_prover_destroy", ctypes.c_int, [_vp, _vp]),
    ("stwo_rocm_prove_poseidon", ctypes.c_int, [_vp, ctypes.POINTER(PoseidonParams), ctypes.c_char_p,
                                                 ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                                 ctypes.POINTER(ProveStats)]),
    ("stwo_rocm_prove_wide_fibonacci", ctypes.c_int, [_vp, ctypes.POINTER(WideFibParams), _vp, ctypes.c_char_p,
                                                       ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                                       ctypes.POINTER(ProveStats)]),
]

_SIGNATURES += [
    ("stwo_rocm_prove_logup_xor", ctypes.c_int, [_vp, ctypes.POINTER(LogupParams), ctypes.c_char_p, ctypes.c_uint64,
                                                  ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_xor_trace", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    ("stwo_rocm_logup_denominators", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_uint64, _u32p, _u32p, _vp]),
    ("stwo_rocm_logup_cumulative", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint64, _vp, _u32p]),
    ("stwo_rocm_logup_cumulative_noalloc", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint64, _vp,
                                                           _u32p]),
    ("stwo_rocm_gkr_gen_eq_evals", ctypes.c_int, [_vp, _u32p, ctypes.c_uint32, _u32p, _vp]),
    ("stwo_rocm_gkr_next_layer", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, ctypes.c_uint32, _vp, _vp]),
    ("stwo_rocm_gkr_sum_as_poly_in_first_variable", ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp, _vp, ctypes.c_uint32,
                                                                    _u32p, _u32p, _u32p, ctypes.c_uint32, _u32p, _u32p,
                                                                    _u32p]),
    ("stwo_rocm_mle_fix_first_variable", ctypes.c_int, [_vp, ctypes.c_int, _vp, ctypes.c_uint32, _u32p, _vp]),
    ("stwo_rocm_gkr_prove_batch", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int),
                                                  ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32p, ctypes.c_char_p,
                                                  _u32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_logup_xor_pipeline", ctypes.c_int, [_vp, ctypes.c_uint32, ctypes.c_uint64, _u32p, ctypes.c_uint64,
                                                     ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_batched_create", ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.POINTER(_vp)]),
    ("stwo_rocm_batched_set_tuning", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_uint32]),
    ("stwo_rocm_batched_autotune", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, _u32p, ctypes.c_char_p,
                                                   ctypes.c_size_t]),
    ("stwo_rocm_batched_run", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_double), _u32p, ctypes.c_uint64,
                                              ctypes.POINTER(ctypes.c_uint64)]),
    ("stwo_rocm_batched_destroy", ctypes.c_int, [_vp, _vp]),
    ("stwo_rocm_serve", ctypes.c_int, [_vp, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    # resident hook surface
    ("stwo_rocm_channel_create", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_uint32, ctypes.POINTER(_vp)]),
    ("stwo_rocm_channel_destroy", ctypes.c_int, [_vp, _vp]),
    ("stwo_rocm_channel_state", ctypes.c_int, [_vp, _vp, ctypes.c_char_p, _u32p]),
    ("stwo_rocm_channel_mix_u32s", ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint64]),
    ("stwo_rocm_channel_mix_u64", ctypes.c_int, [_vp, _vp, ctypes.c_uint64]),
    ("stwo_rocm_channel_mix_felts", ctypes.c_int, [_vp, _vp, _u32p, ctypes.c_uint64]),
    ("stwo_rocm_channel_mix_root", ctypes.c_int, [_vp, _vp, ctypes.c_char_p]),
    ("stwo_rocm_channel_draw_secure_felt", ctypes.c_int, [_vp, _vp, _u32p]),
    ("stwo_rocm_channel_draw_u32s", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p]),
    ("stwo_rocm_prepare_and_commit_owned", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_uint32,
                                                           ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                           ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_prepare_and_commit_polys", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_uint32,
                                                           ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                                           ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]),
    ("stwo_rocm_secure_column_from_line", ctypes.c_int, [_vp, _vp, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    ("stwo_rocm_commit_line", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, ctypes.POINTER(_vp), ctypes.c_char_p]),
    ("stwo_rocm_fri_fold_line_and_commit", ctypes.c_int, [_vp, _vp, ctypes.c_uint32, _u32p, ctypes.c_uint32, _vp,
                                                           ctypes.POINTER(_vp), ctypes.c_char_p]),
    ("stwo_rocm_fri_commit_layers", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint32, ctypes.c_uint32,
                                                    ctypes.POINTER(_vp), ctypes.POINTER(_vp), _vp]),
    ("stwo_rocm_gather_values", ctypes.c_int, [_vp, ctypes.POINTER(_vp), _u32p, ctypes.c_uint32, ctypes.c_uint32,
                                                ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint32,
                                                _u32p]),
    ("stwo_rocm_profile_enable", ctypes.c_int, [_vp, ctypes.c_int]),
    ("stwo_rocm_profile_select", ctypes.c_int, [_vp, ctypes.c_char_p]),
    ("stwo_rocm_profile_sample", ctypes.c_int, [_vp, ctypes.c_uint32]),
    ("stwo_rocm_profile_read", ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t,
                                               ctypes.POINTER(ctypes.c_size_t)]),
]

SYMBOLS = [s[0] for s in _SIGNATURES]


def load_library(path=LIB_PATH):
    """Load the built HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RocmError(-1, "HIP library not built (%s): run __graft_entry__.build()" % path)
    lib = ctypes.CDLL(path)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _u32(a):
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return a, a.ctypes.data_as(_u32p)


def _ptr_array(ptrs):
    arr = (_vp * max(len(ptrs), 1))(*[int(p) for p in ptrs])
    return arr


class DeviceBuffer:
    """A buffer of the context's HBM arena (freed with ``free`` or context close)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _vp()
        ctx._check(ctx.lib.stwo_rocm_buffer_alloc(ctx.h, self.nbytes, ctypes.byref(p)))
        self.ptr = p.value or 0

    def __int__(self):
        return self.ptr

    def offset(self, words):
        return self.ptr + 4 * int(words)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        self.ctx._check(self.ctx.lib.stwo_rocm_buffer_upload(self.ctx.h, self.ptr, arr.ctypes.data, arr.nbytes))
        return self

    def download(self, dtype=np.uint32, count=None, offset_bytes=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset_bytes) // dtype.itemsize
        out = np.empty(count, dtype=dtype)
        self.ctx._check(self.ctx.lib.stwo_rocm_buffer_download(self.ctx.h, out.ctypes.data, self.ptr + offset_bytes,
                                                               out.nbytes))
        return out

    def free(self):
        if self.ptr:
            self.ctx._check(self.ctx.lib.stwo_rocm_buffer_free(self.ctx.h, self.ptr))
            self.ptr = 0


class Tree:
    """A resident lifted Merkle tree (MerkleTree(H): root / maxLogSize / readHashes / decommit)."""

    def __init__(self, ctx, handle, root):
        self.ctx, self.h, self.root = ctx, handle, root

    def max_log_size(self):
        out = ctypes.c_uint32()
        self.ctx._check(self.ctx.lib.stwo_rocm_tree_max_log(self.h, ctypes.byref(out)))
        return out.value

    def read_hashes(self, layer_log, indices):
        idx, pidx = _u32(indices)
        buf = ctypes.create_string_buffer(32 * max(len(idx), 1))
        self.ctx._check(self.ctx.lib.stwo_rocm_tree_read_hashes(self.ctx.h, self.h, layer_log, pidx, len(idx), buf))
        return [buf.raw[32 * i: 32 * i + 32] for i in range(len(idx))]

    def decommit_witness(self, positions):
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        cap = 64 * (len(pos) + 1) * (self.max_log_size() + 1)
        buf = ctypes.create_string_buffer(32 * cap)
        n = ctypes.c_uint32()
        self.ctx._check(self.ctx.lib.stwo_rocm_tree_decommit(self.ctx.h, self.h,
                                                             pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                             len(pos), buf, cap, ctypes.byref(n)))
        return [buf.raw[32 * i: 32 * i + 32] for i in range(n.value)]

    def free(self):
        if self.h:
            self.ctx._check(self.ctx.lib.stwo_rocm_tree_free(self.ctx.h, self.h))
            self.h = None


class DeviceChannel:
    """A Blake2s Fiat-Shamir channel resident in HBM (Blake2sChannel: mix*/draw* run on the device)."""

    def __init__(self, ctx, digest=None, n_draws=0):
        self.ctx = ctx
        h = _vp()
        ctx._check(ctx.lib.stwo_rocm_channel_create(ctx.h, bytes(digest) if digest is not None else None, n_draws,
                                                    ctypes.byref(h)))
        self.h = h.value

    def state(self):
        d = ctypes.create_string_buffer(32)
        n = ctypes.c_uint32()
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_state(self.ctx.h, self.h, d, ctypes.byref(n)))
        return d.raw, n.value

    def mix_u32s(self, words):
        w, pw = _u32(words if len(words) else [0])
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_mix_u32s(self.ctx.h, self.h, pw, len(words)))

    def mix_u64(self, value):
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_mix_u64(self.ctx.h, self.h, int(value)))

    def mix_felts(self, felts):
        f = np.ascontiguousarray(np.asarray(felts, dtype=np.uint32).reshape(-1))
        n = len(f) // 4
        f, pf = _u32(f if len(f) else [0])
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_mix_felts(self.ctx.h, self.h, pf, n))

    def mix_root(self, root):
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_mix_root(self.ctx.h, self.h, bytes(root)))

    def draw_secure_felt(self):
        o, po = _u32(np.zeros(4))
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_draw_secure_felt(self.ctx.h, self.h, po))
        return [int(x) for x in o]

    def draw_u32s(self, n_blocks=1):
        o, po = _u32(np.zeros(8 * max(n_blocks, 1)))
        self.ctx._check(self.ctx.lib.stwo_rocm_channel_draw_u32s(self.ctx.h, self.h, n_blocks, po))
        return [int(x) for x in o[: 8 * n_blocks]]

    def close(self):
        if self.h:
            self.ctx._check(self.ctx.lib.stwo_rocm_channel_destroy(self.ctx.h, self.h))
            self.h = None


class Context:
    """One HIP device context (stream, HBM arena, twiddle tower)."""

    def __init__(self, device=0, hbm_budget_bytes=0):
        self.lib = load_library()
        h = _vp()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.stwo_rocm_ctx_create(device, hbm_budget_bytes, ctypes.byref(h), err, 512)
        if rc != 0:
            raise RocmError(rc, err.value.decode(errors="replace"))
        self.h = h.value

    # ---- lifecycle
    def _check(self, rc):
        if rc != 0:
            raise RocmError(rc, self.lib.stwo_rocm_last_error(self.h).decode(errors="replace"))

    def close(self):
        if self.h:
            self.lib.stwo_rocm_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def identity(self):
        buf = ctypes.create_string_buffer(512)
        self._check(self.lib.stwo_rocm_ctx_identity(self.h, buf, 512))
        return buf.value.decode()

    def synchronize(self):
        self._check(self.lib.stwo_rocm_synchronize(self.h))

    def memory_info(self):
        a, b, c = ctypes.c_uint64(), ctypes.c_uint64(), ctypes.c_uint64()
        self._check(self.lib.stwo_rocm_memory_info(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"live": a.value, "cached": b.value, "peak": c.value}

    def release_cached(self):
        self._check(self.lib.stwo_rocm_release_cached(self.h))

    def set_budget(self, hbm_budget_bytes):
        """Device budget for later allocations (0 = unlimited); cached blocks beyond it are released."""
        self._check(self.lib.stwo_rocm_set_budget(self.h, int(hbm_budget_bytes)))

    def runtime_info(self, reset_peak=False):
        """Twiddle tower (circle log, bytes, builds) and arena counters; reset_peak restarts the peak."""
        info = RuntimeInfo()
        self._check(self.lib.stwo_rocm_get_runtime_info(self.h, 1 if reset_peak else 0, ctypes.byref(info)))
        return {k: getattr(info, k) for k, _ in RuntimeInfo._fields_ if k != "reserved"}

    # ---- per-kernel device-time profile (HIP events on the context stream)
    def profile_enable(self, enable=True):
        self._check(self.lib.stwo_rocm_profile_enable(self.h, 1 if enable else 0))

    def profile_select(self, kernel_class=None):
        """Record only launches of `kernel_class` (a class name or a list of names; None: every class)."""
        if isinstance(kernel_class, (list, tuple)):
            kernel_class = "|".join(kernel_class)
        self._check(self.lib.stwo_rocm_profile_select(self.h, kernel_class.encode() if kernel_class else None))

    def profile_sample(self, every=1):
        """Record only every `every`-th launch of the selected class."""
        self._check(self.lib.stwo_rocm_profile_sample(self.h, every))

    def profile_read(self):
        import json

        cap = 1 << 16
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_size_t()
        self._check(self.lib.stwo_rocm_profile_read(self.h, buf, cap, ctypes.byref(n)))
        return json.loads(buf.value.decode())

    # ---- buffers
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr):
        arr = np.ascontiguousarray(arr)
        return DeviceBuffer(self, max(arr.nbytes, 4)).upload(arr)

    # ---- twiddles (M31TwiddleTower)
    def prepare_twiddles(self, max_circle_log):
        self._check(self.lib.stwo_rocm_twiddles_prepare(self.h, max_circle_log))

    def read_twiddles(self, circle_log):
        n = 1 << (circle_log - 1)
        tw, ptw = _u32(np.zeros(n))
        itw, pitw = _u32(np.zeros(n))
        self._check(self.lib.stwo_rocm_twiddles_read(self.h, circle_log, ptw, pitw))
        return tw, itw

    # ---- FieldOps / AccumulationOps
    def batch_inverse(self, d_in, n, d_out):
        self._check(self.lib.stwo_rocm_batch_inverse(self.h, int(d_in), n, int(d_out)))

    def bit_reverse(self, d_vals, log_size, words_per_elem=1):
        self._check(self.lib.stwo_rocm_bit_reverse(self.h, int(d_vals), log_size, words_per_elem))

    def accumulate(self, d_acc, d_src, n_words):
        self._check(self.lib.stwo_rocm_accumulate(self.h, int(d_acc), int(d_src), n_words))

    # ---- PolyOps: interpolateCircleBuffers / evaluateCircleBuffers / evalAtPoint
    def interpolate_circle_buffers(self, cols, log_size):
        arr = _ptr_array(cols)
        ms = ctypes.c_double()
        self._check(self.lib.stwo_rocm_circle_interpolate(self.h, arr, len(cols), log_size, ctypes.byref(ms)))
        return ms.value

    def evaluate_circle_buffers(self, coeffs, log_size, log_blowup, out):
        a, b = _ptr_array(coeffs), _ptr_array(out)
        ms = ctypes.c_double()
        self._check(self.lib.stwo_rocm_circle_evaluate(self.h, a, len(coeffs), log_size, log_blowup, b,
                                                       ctypes.byref(ms)))
        return ms.value

    def eval_at_point(self, coeffs, log_size, point8):
        a = _ptr_array(coeffs)
        p, pp = _u32(point8)
        o, po = _u32(np.zeros(4 * max(len(coeffs), 1)))
        self._check(self.lib.stwo_rocm_eval_at_point(self.h, a, len(coeffs), log_size, pp, po))
        return o[: 4 * len(coeffs)].reshape(-1, 4)

    # ---- MerkleOps
    def commit_merkle(self, cols, log_sizes):
        a = _ptr_array(cols)
        logs, plogs = _u32(log_sizes if len(log_sizes) else [0])
        h = _vp()
        root = ctypes.create_string_buffer(32)
        ms = ctypes.c_double()
        self._check(self.lib.stwo_rocm_merkle_commit(self.h, a, plogs, len(cols), ctypes.byref(h), root,
                                                     ctypes.byref(ms)))
        return Tree(self, h.value, root.raw)

    def merkle_parent_layer(self, d_children, log_size, d_parents):
        self._check(self.lib.stwo_rocm_merkle_parent_layer(self.h, int(d_children), log_size, int(d_parents)))

    # ---- FriOps
    def fold_line(self, d_src, log_src, alpha, d_dst):
        a, pa = _u32(alpha)
        self._check(self.lib.stwo_rocm_fri_fold_line(self.h, int(d_src), log_src, pa, int(d_dst)))

    def fold_circle_into_line(self, coords, log_src, alpha, d_dst):
        a, pa = _u32(alpha)
        self._check(self.lib.stwo_rocm_fri_fold_circle_into_line(self.h, _ptr_array(coords), log_src, pa,
                                                                 int(d_dst)))

    # ---- resident hooks (MetalCommitBackend hook set)
    def channel(self, digest=None, n_draws=0):
        return DeviceChannel(self, digest, n_draws)

    def _prepare_and_commit(self, fn, cols, log_sizes, log_blowup, ext_out):
        logs, plogs = _u32(log_sizes if len(log_sizes) else [0])
        h = _vp()
        root = ctypes.create_string_buffer(32)
        ms = ctypes.c_double()
        self._check(fn(self.h, _ptr_array(cols), plogs, len(cols), log_blowup, _ptr_array(ext_out), ctypes.byref(h),
                       root, ctypes.byref(ms)))
        return Tree(self, h.value, root.raw)

    def prepare_and_commit_owned(self, cols, log_sizes, log_blowup, ext_out):
        """prepareAndCommitOwned: evaluations -> coefficients (in place) -> LDE (ext_out) -> tree."""
        return self._prepare_and_commit(self.lib.stwo_rocm_prepare_and_commit_owned, cols, log_sizes, log_blowup,
                                        ext_out)

    def prepare_and_commit_polys(self, coeffs, log_sizes, log_blowup, ext_out):
        """prepareAndCommitPolys: coefficients -> LDE (ext_out) -> tree."""
        return self._prepare_and_commit(self.lib.stwo_rocm_prepare_and_commit_polys, coeffs, log_sizes, log_blowup,
                                        ext_out)

    def secure_column_from_line(self, d_line, n, coords):
        self._check(self.lib.stwo_rocm_secure_column_from_line(self.h, int(d_line), n, _ptr_array(coords)))

    def commit_line(self, d_line, log_size):
        h = _vp()
        root = ctypes.create_string_buffer(32)
        self._check(self.lib.stwo_rocm_commit_line(self.h, int(d_line), log_size, ctypes.byref(h), root))
        return Tree(self, h.value, root.raw)

    def fold_line_and_commit_next(self, d_src, log_src, alpha, n_folds, d_dst):
        a, pa = _u32(alpha)
        h = _vp()
        root = ctypes.create_string_buffer(32)
        self._check(self.lib.stwo_rocm_fri_fold_line_and_commit(self.h, int(d_src), log_src, pa, n_folds, int(d_dst),
                                                                ctypes.byref(h), root))
        return Tree(self, h.value, root.raw)

    def commit_fri_layers(self, channel, d_line, log_size, log_last_size, d_lines, d_last_line):
        """commitFriLayers: trees of the lines of log log_size .. log_last_size+1 (channel on the device)."""
        n = log_size - log_last_size
        trees = (_vp * n)()
        self._check(self.lib.stwo_rocm_fri_commit_layers(self.h, channel.h, int(d_line), log_size, log_last_size,
                                                         _ptr_array(d_lines), trees, int(d_last_line)))
        out = []
        for k in range(n):
            t = Tree(self, trees[k], None)
            t.root = self.tree_root(t)
            out.append(t)
        return out

    def tree_root(self, tree):
        root = ctypes.create_string_buffer(32)
        self._check(self.lib.stwo_rocm_tree_root(self.h, tree.h, root))
        return root.raw

    def gather_values(self, cols, log_sizes, lifting_log, positions, words_per_value=1):
        pos = np.ascontiguousarray(positions, dtype=np.uint64)
        logs, plogs = _u32(log_sizes if len(log_sizes) else [0])
        o, po = _u32(np.zeros(max(len(cols) * len(pos) * words_per_value, 1)))
        self._check(self.lib.stwo_rocm_gather_values(self.h, _ptr_array(cols), plogs, len(cols), words_per_value,
                                                     lifting_log, pos.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                     len(pos), po))
        return o[: len(cols) * len(pos) * words_per_value].reshape(len(cols), len(pos), words_per_value)

    # ---- quotients (computeLazyQuotients / commitLazyMerkle)
    def compute_quotients(self, lifting_log_size, columns, column_log_sizes, sample_columns, sample_points,
                          sample_values, random_coeff, out_coords, commit=False):
        logs, plogs = _u32(column_log_sizes if len(column_log_sizes) else [0])
        sc, psc = _u32(sample_columns if len(sample_columns) else [0])
        sp, psp = _u32(np.asarray(sample_points, dtype=np.uint32).reshape(-1) if len(sample_points) else [0])
        sv, psv = _u32(np.asarray(sample_values, dtype=np.uint32).reshape(-1) if len(sample_values) else [0])
        rc, prc = _u32(random_coeff)
        h = _vp()
        self._check(self.lib.stwo_rocm_compute_quotients(self.h, lifting_log_size, _ptr_array(columns), plogs,
                                                         len(columns), psc, psp, psv, len(sample_columns), prc,
                                                         _ptr_array(out_coords),
                                                         ctypes.byref(h) if commit else None))
        if commit:
            root = ctypes.create_string_buffer(32)
            self._check(self.lib.stwo_rocm_tree_root(self.h, h.value, root))
            return Tree(self, h.value, root.raw)
        return None

    # ---- wide Fibonacci
    def fill_quadratic_recurrence_trace(self, log_n_rows, seq_len, d_out, mode=0, seed=0):
        self._check(self.lib.stwo_rocm_widefib_trace(self.h, log_n_rows, seq_len, mode, seed, int(d_out)))

    def compute_composition_evaluation(self, eval_cols, log_n_rows, random_coeff, out_coords):
        rc, prc = _u32(random_coeff)
        self._check(self.lib.stwo_rocm_widefib_composition(self.h, _ptr_array(eval_cols), len(eval_cols), log_n_rows,
                                                           prc, _ptr_array(out_coords)))

    # ---- channel
    def grind(self, digest, pow_bits):
        out = ctypes.c_uint64()
        self._check(self.lib.stwo_rocm_pow_grind(self.h, bytes(digest), pow_bits, ctypes.byref(out)))
        return out.value

    # ---- whole proof
    def prove_wide_fibonacci(self, log_n_rows, sequence_len, pow_bits=10, log_blowup_factor=1,
                             log_last_layer_degree_bound=0, n_queries=3, trace_mode=0, trace_seed=0, d_trace=None,
                             stats=True):
        """Returns (proof bytes, stage timings dict or None).  stats=False records no stage
        events on the stream (each costs a few microseconds of GPU time)."""
        prm = WideFibParams(log_n_rows, sequence_len, pow_bits, log_blowup_factor, log_last_layer_degree_bound,
                            n_queries, trace_mode, trace_seed)
        st = ProveStats() if stats else None
        cap = 1 << 20
        n = ctypes.c_uint64()
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self.lib.stwo_rocm_prove_wide_fibonacci(self.h, ctypes.byref(prm),
                                                         int(d_trace) if d_trace is not None else None, buf, cap,
                                                         ctypes.byref(n), ctypes.byref(st) if st is not None else None)
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return buf.raw[: n.value], (st.as_dict() if st is not None else None)


    def prove_wide_fibonacci_batch(self, log_n_rows, sequence_len, seeds=None, n_proofs=None, pow_bits=10,
                                   log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3, trace_mode=1,
                                   trace_seed=0):
        """Many wide-Fibonacci proofs of one shape in single launches: proof k uses seeds[k] (or
        trace_seed for all n_proofs when seeds is None).  Returns (list of proof bytes, stats dict)."""
        prm = WideFibParams(log_n_rows, sequence_len, pow_bits, log_blowup_factor, log_last_layer_degree_bound,
                            n_queries, trace_mode, trace_seed)
        arr, n_proofs = _seed_array(seeds, n_proofs)
        return self._batch_call(
            lambda buf, cap, offs, st: self.lib.stwo_rocm_prove_wide_fibonacci_batch(
                self.h, ctypes.byref(prm), n_proofs, arr, buf, cap, offs, ctypes.byref(st)),
            n_proofs, log_n_rows + log_blowup_factor, sequence_len, n_queries, log_last_layer_degree_bound)

    def batch_prover(self, log_n_rows, sequence_len, n_proofs, pow_bits=10, log_blowup_factor=1,
                     log_last_layer_degree_bound=0, n_queries=3, trace_mode=1, trace_seed=0):
        """A reusable plan for batches of n_proofs proofs of one shape (BatchProver.run)."""
        return BatchProver(self, log_n_rows, sequence_len, n_proofs, pow_bits, log_blowup_factor,
                           log_last_layer_degree_bound, n_queries, trace_mode, trace_seed)

    def _batch_call(self, call, n_proofs, L, sequence_len, n_queries, last_bound):
        offs = (ctypes.c_uint64 * (n_proofs + 1))()
        st = BatchStats()
        # an upper bound of one proof's wire size (sampled values, commitments, hash witnesses
        # of every tree and FRI layer, queried values, the last layer), kept in one buffer reused
        # across calls
        per = (8192 + 48 * (sequence_len + 8) + 12 * n_queries * (sequence_len + 8) +
               n_queries * (L + 1) * (4 * 132 + 48) + 48 * (1 << last_bound))
        cap = max(n_proofs * per, getattr(self, "_batch_cap", 0))
        while True:
            if getattr(self, "_batch_cap", 0) < cap:
                self._batch_buf = ctypes.create_string_buffer(cap)
                self._batch_cap = cap
            buf, cap = self._batch_buf, self._batch_cap
            rc = call(buf, cap, offs, st)
            if rc == 1 and offs[n_proofs] > cap:
                cap = offs[n_proofs]
                continue
            self._check(rc)
            addr = ctypes.addressof(buf)
            o = list(offs)
            proofs = [ctypes.string_at(addr + o[k], o[k + 1] - o[k]) for k in range(n_proofs)]
            phases = ("setup", "device_wait", "last_layers", "pow", "queries_gather", "encode")
            return proofs, {"device_ms": st.device_ms, "tail_ms": st.tail_ms, "total_ms": st.total_ms,
                            "slice_bytes": st.slice_bytes,
                            "phase_ms": {k: round(st.phase_ms[i], 3) for i, k in enumerate(phases)}}

    def prove_poseidon(self, log_n_instances, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0,
                       n_queries=3, stats=True):
        """Poseidon AIR proof (examples/poseidon.zig:67-73): (proof bytes, stage timings or None)."""
        prm = PoseidonParams(log_n_instances, pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries)
        st = ProveStats() if stats else None
        cap = 1 << 20
        n = ctypes.c_uint64()
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self.lib.stwo_rocm_prove_poseidon(self.h, ctypes.byref(prm), buf, cap, ctypes.byref(n),
                                                   ctypes.byref(st) if st is not None else None)
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return buf.raw[: n.value], (st.as_dict() if st is not None else None)

    def prove_logup_xor(self, log_n, seed, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0,
                        n_queries=3):
        """XOR-lookup AIR proof (the LogUp benchmark's AIR as a whole proof): proof bytes."""
        prm = LogupParams(log_n, pow_bits, seed, log_blowup_factor, log_last_layer_degree_bound, n_queries, 0)
        cap = 1 << 20
        n = ctypes.c_uint64()
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self.lib.stwo_rocm_prove_logup_xor(self.h, ctypes.byref(prm), buf, cap, ctypes.byref(n))
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return buf.raw[: n.value]

    def prove_xor(self, log_size, log_step, offset, pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0,
                  n_queries=3, stats=True):
        """Periodic XOR example AIR proof (examples/xor.zig): (proof bytes, stage timings or None)."""
        prm = XorParams(log_size, log_step, offset, pow_bits, log_blowup_factor, log_last_layer_degree_bound,
                        n_queries)
        st = ProveStats() if stats else None
        cap = 1 << 20
        n = ctypes.c_uint64()
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = self.lib.stwo_rocm_prove_xor(self.h, ctypes.byref(prm), buf, cap, ctypes.byref(n),
                                              ctypes.byref(st) if st is not None else None)
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return buf.raw[: n.value], (st.as_dict() if st is not None else None)

    # ---- LogUp / GKR (kernels_logup.hip, gkr.hip)
    def xor_trace(self, log_size, seed, d_a, d_b, d_c, d_mult):
        self._check(self.lib.stwo_rocm_xor_trace(self.h, log_size, seed, int(d_a), int(d_b), int(d_c), int(d_mult)))

    def logup_denominators(self, d_a, d_b, d_c, n, z, alpha, d_den):
        zz, pz = _u32(z)
        al, pal = _u32(alpha)
        self._check(self.lib.stwo_rocm_logup_denominators(
            self.h, int(d_a) if d_a is not None else None, int(d_b) if d_b is not None else None,
            int(d_c) if d_c is not None else None, n, pz, pal, int(d_den)))

    def logup_cumulative(self, num_kind, d_num, d_den, n, d_sums):
        cl, pcl = _u32(np.zeros(4))
        self._check(self.lib.stwo_rocm_logup_cumulative(self.h, num_kind, int(d_num) if d_num is not None else None,
                                                        int(d_den), n, int(d_sums), pcl))
        return cl

    def logup_cumulative_noalloc(self, num_kind, d_num, d_den, n, d_sums):
        """logup_cumulative with no allocation (fixed per-device scratch array; n <= 2^27)."""
        cl, pcl = _u32(np.zeros(4))
        self._check(self.lib.stwo_rocm_logup_cumulative_noalloc(self.h, num_kind,
                                                                int(d_num) if d_num is not None else None,
                                                                int(d_den), n, int(d_sums), pcl))
        return cl

    def gkr_gen_eq_evals(self, y, scale, d_out):
        y = np.ascontiguousarray(y, dtype=np.uint32)
        yy, py = _u32(y if y.size else np.zeros(4))
        sc, psc = _u32(scale)
        self._check(self.lib.stwo_rocm_gkr_gen_eq_evals(self.h, py, y.size // 4, psc, int(d_out)))

    def gkr_next_layer(self, kind, d_num, d_den, n_vars, d_out_num, d_out_den):
        self._check(self.lib.stwo_rocm_gkr_next_layer(self.h, kind, int(d_num) if d_num is not None else None,
                                                      int(d_den), n_vars,
                                                      int(d_out_num) if d_out_num is not None else None,
                                                      int(d_out_den)))

    def gkr_sum_as_poly_in_first_variable(self, kind, d_eq, d_num, d_den, n_vars, lam, claim, y, correction):
        la, pla = _u32(lam)
        cl, pcl = _u32(claim)
        yy, py = _u32(y)
        co, pco = _u32(correction)
        out, pout = _u32(np.zeros(16))
        n, pn = _u32(np.zeros(1))
        self._check(self.lib.stwo_rocm_gkr_sum_as_poly_in_first_variable(
            self.h, kind, int(d_eq), int(d_num) if d_num is not None else None, int(d_den), n_vars, pla, pcl, py,
            yy.size // 4, pco, pout, pn))
        return out[: 4 * int(n[0])]

    def mle_fix_first_variable(self, base_field, d_in, n_vars, r, d_out):
        rr, pr = _u32(r)
        self._check(self.lib.stwo_rocm_mle_fix_first_variable(self.h, 1 if base_field else 0, int(d_in), n_vars, pr,
                                                              int(d_out)))

    def gkr_prove_batch(self, instances, digest=bytes(32)):
        """instances: list of (kind, d_num or None, d_den, n_vars).  Returns the proof word stream."""
        k = len(instances)
        kinds = (ctypes.c_int * k)(*[i[0] for i in instances])
        nums = (_vp * k)(*[int(i[1]) if i[1] is not None else None for i in instances])
        dens = (_vp * k)(*[int(i[2]) for i in instances])
        nv, pnv = _u32([i[3] for i in instances])
        cap = 1 << 16
        n = ctypes.c_uint64()
        while True:
            out, pout = _u32(np.zeros(cap))
            rc = self.lib.stwo_rocm_gkr_prove_batch(self.h, k, kinds, nums, dens, pnv, bytes(digest), pout, cap,
                                                    ctypes.byref(n))
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return out[: n.value].copy()


    def logup_xor_pipeline(self, log_n, seed):
        """LogUp workload word stream and the call's host wall time (ms)."""
        cap = 1 << 16
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        while True:
            out, pout = _u32(np.zeros(cap))
            rc = self.lib.stwo_rocm_logup_xor_pipeline(self.h, log_n, seed, pout, cap, ctypes.byref(n),
                                                       ctypes.byref(ms))
            if rc == 1 and n.value > cap:
                cap = n.value
                continue
            self._check(rc)
            return out[: n.value].copy(), ms.value


    def batched(self, n_columns=64, log_min=14, log_max=22, log_blowup=2, fri_log=22, n_folds=3, seed=0):
        return BatchedPipeline(self, n_columns, log_min, log_max, log_blowup, fri_log, n_folds, seed)


class BatchedSpec(ctypes.Structure):
    _fields_ = [("n_columns", ctypes.c_uint32), ("log_min", ctypes.c_uint32), ("log_max", ctypes.c_uint32),
                ("log_blowup", ctypes.c_uint32), ("fri_log", ctypes.c_uint32), ("n_folds", ctypes.c_uint32),
                ("seed", ctypes.c_uint64)]


class BatchedPipeline:
    """Planned batched LDE + Merkle + FRI pipeline replayed as one hipGraph."""

    def __init__(self, ctx, *spec):
        self.ctx = ctx
        self.spec = BatchedSpec(*spec)
        h = _vp()
        ctx._check(ctx.lib.stwo_rocm_batched_create(ctx.h, ctypes.addressof(self.spec), ctypes.byref(h)))
        self.h = h

    def set_tuning(self, fft_tile_log, merkle_levels):
        self.ctx._check(self.ctx.lib.stwo_rocm_batched_set_tuning(self.ctx.h, self.h, fft_tile_log, merkle_levels))

    def autotune(self, reps=3):
        import json
        t, pt = _u32(np.zeros(1))
        l, pl = _u32(np.zeros(1))
        buf = ctypes.create_string_buffer(1 << 14)
        self.ctx._check(self.ctx.lib.stwo_rocm_batched_autotune(self.ctx.h, self.h, reps, pt, pl, buf, 1 << 14))
        return int(t[0]), int(l[0]), json.loads(buf.value.decode())

    def run(self, n_runs=1, graph=True):
        out, pout = _u32(np.zeros(4096))
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        self.ctx._check(self.ctx.lib.stwo_rocm_batched_run(self.ctx.h, self.h, n_runs, 1 if graph else 0,
                                                           ctypes.byref(ms), pout, 4096, ctypes.byref(n)))
        return out[: n.value].copy(), ms.value

    def close(self):
        if self.h:
            self.ctx._check(self.ctx.lib.stwo_rocm_batched_destroy(self.ctx.h, self.h))
            self.h = None


def _seed_array(seeds, n_proofs):
    if seeds is None:
        return None, int(n_proofs)
    seeds = [int(x) for x in seeds]
    return (ctypes.c_uint64 * len(seeds))(*seeds), len(seeds)


class BatchProver:
    """stwo_rocm_batch_prover_*: the device allocation, tables and pinned readback of a batch
    shape, set up once; run(seeds) proves one batch (graph=True records the device part of the
    first run into a hipGraph and replays it afterwards)."""

    def __init__(self, ctx, log_n_rows, sequence_len, n_proofs, pow_bits, log_blowup_factor,
                 log_last_layer_degree_bound, n_queries, trace_mode, trace_seed):
        self.ctx, self.n = ctx, int(n_proofs)
        self.shape = (log_n_rows + log_blowup_factor, sequence_len, n_queries, log_last_layer_degree_bound)
        prm = WideFibParams(log_n_rows, sequence_len, pow_bits, log_blowup_factor, log_last_layer_degree_bound,
                            n_queries, trace_mode, trace_seed)
        h = _vp()
        ctx._check(ctx.lib.stwo_rocm_batch_prover_create(ctx.h, ctypes.byref(prm), self.n, ctypes.byref(h)))
        self.h = h

    def run(self, seeds=None, graph=True):
        arr, n = _seed_array(seeds, self.n)
        if n != self.n:
            raise ValueError("the plan proves %d proofs per run" % self.n)
        L, W, q, lb = self.shape
        return self.ctx._batch_call(
            lambda buf, cap, offs, st: self.ctx.lib.stwo_rocm_batch_prover_run(
                self.ctx.h, self.h, arr, 1 if graph else 0, buf, cap, offs, ctypes.byref(st)),
            self.n, L, W, q, lb)

    def close(self):
        if self.h:
            self.ctx._check(self.ctx.lib.stwo_rocm_batch_prover_destroy(self.ctx.h, self.h))
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class BatchStats(ctypes.Structure):
    _fields_ = [("device_ms", ctypes.c_double), ("tail_ms", ctypes.c_double), ("total_ms", ctypes.c_double),
                ("slice_bytes", ctypes.c_uint64), ("phase_ms", ctypes.c_double * 6)]


class RuntimeInfo(ctypes.Structure):
    _fields_ = [("tower_log", ctypes.c_uint32), ("reserved", ctypes.c_uint32), ("tower_bytes", ctypes.c_uint64),
                ("tower_builds", ctypes.c_uint64), ("bytes_live", ctypes.c_uint64), ("bytes_cached", ctypes.c_uint64),
                ("peak_live", ctypes.c_uint64)]


def job_device_bytes(kind, log_n, seq_len=0):
    """Device-byte estimate stwo_rocm_serve admits a job of this kind and shape against (no device needed)."""
    lib = load_library()
    out = ctypes.c_uint64()
    rc = lib.stwo_rocm_job_device_bytes(ctypes.byref(ServeJob(kind, log_n, seq_len, 0, 0, 0.0)), ctypes.byref(out))
    if rc != 0:
        raise RocmError("job_device_bytes: invalid job shape")
    return out.value


class ServeJob(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("log_n", ctypes.c_uint32), ("seq_len", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("seed", ctypes.c_uint64), ("arrival_ms", ctypes.c_double)]


class ServeResult(ctypes.Structure):
    _fields_ = [("digest", ctypes.c_uint8 * 32), ("arrival_ms", ctypes.c_double), ("start_ms", ctypes.c_double),
                ("end_ms", ctypes.c_double), ("slot", ctypes.c_uint32), ("status", ctypes.c_uint32),
                ("output_bytes", ctypes.c_uint64)]


class ServeStats(ctypes.Structure):
    _fields_ = [("makespan_ms", ctypes.c_double), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("jobs_per_s", ctypes.c_double),
                ("peak_admitted_bytes", ctypes.c_uint64), ("peak_device_bytes", ctypes.c_uint64),
                ("max_in_flight", ctypes.c_uint32),
                ("n_slots", ctypes.c_uint32)]


def serve(ctx, jobs, n_slots, hbm_budget_bytes=0):
    """jobs: list of (kind, log_n, seq_len, seed, arrival_ms).  Returns (results as dicts, stats dict)."""
    arr = (ServeJob * len(jobs))(*[ServeJob(k, n, w, 0, s, a) for k, n, w, s, a in jobs])
    res = (ServeResult * len(jobs))()
    st = ServeStats()
    ctx._check(ctx.lib.stwo_rocm_serve(ctx.h, ctypes.addressof(arr), len(jobs), n_slots, hbm_budget_bytes,
                                       ctypes.addressof(res), ctypes.addressof(st)))
    out = [{"digest": bytes(r.digest).hex(), "arrival_ms": r.arrival_ms, "start_ms": r.start_ms,
            "end_ms": r.end_ms, "slot": r.slot, "status": r.status, "output_bytes": r.output_bytes} for r in res]
    stats = {k: getattr(st, k) for k, _ in ServeStats._fields_}
    return out, stats


# Protocol presets (src/prover/native/config.zig:139-159).
PROTOCOLS = {
    "functional": dict(pow_bits=10, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=3),
    "secure": dict(pow_bits=26, log_blowup_factor=1, log_last_layer_degree_bound=0, n_queries=70),
}
