"""The library's field arithmetic (fields.hpp, compiled for the host) against an independent
big-integer restatement of the field laws (m31.zig, cm31.zig, qm31.zig: CM31 = M31[i]/(i^2+1),
QM31 = CM31[u]/(u^2 - (2 + i))).  The QM31 and CM31 products are written out as raw products
reduced once per coordinate and the inverse is an addition chain, so random and edge operands
(0, 1, p - 1, 2^30, ...) are checked against plain Python integers."""
import os
import random
import shutil
import subprocess

import pytest

P = (1 << 31) - 1
HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def cm_mul(x, y):
    a, b = x
    c, d = y
    return ((a * c - b * d) % P, (a * d + b * c) % P)


def qm_mul(x, y):
    x0, x1 = (x[0], x[1]), (x[2], x[3])
    y0, y1 = (y[0], y[1]), (y[2], y[3])
    r = (2, 1)  # u^2 = 2 + i
    lo = cm_mul(x0, y0)
    hi = cm_mul(cm_mul(x1, y1), r)
    c = cm_mul(x0, y1)
    e = cm_mul(x1, y0)
    return ((lo[0] + hi[0]) % P, (lo[1] + hi[1]) % P, (c[0] + e[0]) % P, (c[1] + e[1]) % P)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if not (os.path.exists(HIPCC) or shutil.which("hipcc")):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("fields") / "fields_host")
    src = os.path.join(HERE, "fields_host", "fields_host.cpp")
    r = subprocess.run([HIPCC, "-O2", "-std=c++17", "-x", "hip", src, "-o", out], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("hipcc failed:\n" + r.stderr[-4000:])
    return out


def run(driver, lines):
    r = subprocess.run([driver], input="\n".join(lines) + "\n", capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [tuple(int(v) for v in l.split()) for l in r.stdout.strip().splitlines()]


EDGE = [0, 1, 2, P - 1, P - 2, 1 << 30, (1 << 30) - 1, 65536]


def operands(rng, n):
    for k in range(n):
        yield tuple(rng.choice(EDGE) if (k + j) % 5 == 0 else rng.randrange(P) for j in range(8))


def test_qm31_and_cm31_products(driver):
    rng = random.Random(2024)
    ops = list(operands(rng, 3000))
    ops += [tuple(EDGE[(k >> (3 * j)) % 8] for j in range(8)) for k in range(0, 8 ** 4, 3)]
    got = run(driver, ["q " + " ".join(map(str, v)) for v in ops])
    assert got == [qm_mul(v[:4], v[4:]) for v in ops]
    got = run(driver, ["c " + " ".join(map(str, v[:4])) for v in ops])
    assert got == [cm_mul(v[:2], v[2:4]) for v in ops]


def test_m31_inverse(driver):
    rng = random.Random(7)
    vals = EDGE[1:] + [rng.randrange(1, P) for _ in range(3000)]
    got = run(driver, ["i %d" % a for a in vals])
    assert [g[0] for g in got] == [pow(a, P - 2, P) for a in vals]
    assert all(g[0] * a % P == 1 for g, a in zip(got, vals))
    assert run(driver, ["i 0"]) == [(0,)]
