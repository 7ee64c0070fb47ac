"""The allocation-free LogUp cumulative column (stwo_rocm_logup_cumulative_noalloc) against the
allocating entry point (stwo_rocm_logup_cumulative), through the C-ABI: two calls in a row on one
device, the second with a smaller input, for every numerator kind; no arena bytes change across
a call; sizes beyond the scratch array and zero denominators are rejected."""
import sys

import numpy as np
import pytest

import stwo_rocm

pytestmark = pytest.mark.gpu
P = stwo_rocm.P


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _inputs(rng, num_kind, n):
    den = rng.integers(1, P, size=4 * n, dtype=np.uint32)
    if num_kind == 0:
        num = rng.integers(0, P, size=4 * n, dtype=np.uint32)
    elif num_kind == 1:
        num = rng.integers(0, P, size=n, dtype=np.uint32)
    else:
        num = None
    return num, den


@pytest.mark.parametrize("num_kind", [0, 1, 2])
def test_noalloc_cumulative_equals_allocating_entry(ctx, num_kind):
    rng = np.random.default_rng(700 + num_kind)
    for n in (3 * (1 << 20) + 123, 5000):  # the second call smaller than the first
        _say("noalloc cumulative", num_kind, n)
        num, den = _inputs(rng, num_kind, n)
        dn = ctx.to_device(num) if num is not None else None
        dd = ctx.to_device(den)
        s_ref = ctx.alloc(16 * n)
        s_got = ctx.alloc(16 * n)
        cl_ref = ctx.logup_cumulative(num_kind, dn, dd, n, s_ref)
        before = ctx.memory_info()
        cl_got = ctx.logup_cumulative_noalloc(num_kind, dn, dd, n, s_got)
        assert ctx.memory_info() == before  # nothing allocated, cached or freed
        assert np.array_equal(cl_got, cl_ref)
        assert np.array_equal(s_got.download(), s_ref.download())
        for b in (dn, dd, s_ref, s_got):
            if b is not None:
                b.free()


def test_noalloc_cumulative_rejects_oversize_and_zero_denominator(ctx):
    one = np.zeros(4 * 64, dtype=np.uint32)
    one[::4] = 1  # every denominator is the QM31 one
    dd = ctx.to_device(one)
    ds = ctx.alloc(16 * 64)
    # more rows than the scratch array covers: refused on the host before any launch
    with pytest.raises(stwo_rocm.RocmError) as e:
        ctx.logup_cumulative_noalloc(2, None, dd, (1 << 27) + 1, ds)
    assert e.value.code == 1 and "2^27" in str(e.value)
    den = one.copy()
    den[4 * 17] = 0
    dz = ctx.to_device(den)
    with pytest.raises(stwo_rocm.RocmError):
        ctx.logup_cumulative_noalloc(2, None, dz, 64, ds)
    # the flag starts from zero again on the next call
    cl = ctx.logup_cumulative_noalloc(2, None, dd, 64, ds)
    assert np.array_equal(cl, [64, 0, 0, 0])
    for b in (dd, ds, dz):
        b.free()
