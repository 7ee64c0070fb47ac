"""Independent pure-Python restatements of the field and circle mathematics,
written from the definitions (not from the oracle), used to pin the oracle and
the HIP kernels on small inputs.

M31 = Z/(2^31-1); CM31 = M31[i]/(i^2+1); QM31 = CM31[u]/(u^2-(2+i));
circle group x^2+y^2=1 with (x1,y1)*(x2,y2) = (x1x2-y1y2, x1y2+y1x2),
generator (2, 1268011823) of order 2^31 (src/core/circle.zig:137-143).
"""
P = (1 << 31) - 1
GEN = (2, 1268011823)


def cm(a, b):
    return (a % P, b % P)


def cmul(x, y):
    return ((x[0] * y[0] - x[1] * y[1]) % P, (x[0] * y[1] + x[1] * y[0]) % P)


def cadd(x, y):
    return ((x[0] + y[0]) % P, (x[1] + y[1]) % P)


def csub(x, y):
    return ((x[0] - y[0]) % P, (x[1] - y[1]) % P)


R = (2, 1)


def qmul(x, y):
    """x, y = (a0, a1, b0, b1) meaning (a0 + a1 i) + (b0 + b1 i) u."""
    a, b = (x[0], x[1]), (x[2], x[3])
    c, d = (y[0], y[1]), (y[2], y[3])
    re = cadd(cmul(a, c), cmul(R, cmul(b, d)))
    im = cadd(cmul(a, d), cmul(b, c))
    return (re[0], re[1], im[0], im[1])


def qadd(x, y):
    return tuple((a + b) % P for a, b in zip(x, y))


def qsub(x, y):
    return tuple((a - b) % P for a, b in zip(x, y))


def qpow(x, e):
    r = (1, 0, 0, 0)
    while e:
        if e & 1:
            r = qmul(r, x)
        x = qmul(x, x)
        e >>= 1
    return r


def qinv(x):
    return qpow(x, P ** 4 - 2)


def padd(p, q):
    return ((p[0] * q[0] - p[1] * q[1]) % P, (p[0] * q[1] + p[1] * q[0]) % P)


def pmul(p, k):
    r = (1, 0)
    while k:
        if k & 1:
            r = padd(r, p)
        p = padd(p, p)
        k >>= 1
    return r


def point_at_index(i):
    return pmul(GEN, i % (1 << 31))


def domain_index(log_size, i):
    """CanonicCoset(log).circleDomain(): half coset (initial 2^(31-log-1), step 2^(31-log+1)) then its conjugate."""
    half = 1 << (log_size - 1)
    init, step = 1 << (31 - log_size - 1), 1 << (31 - log_size + 1)
    if i < half:
        return (init + step * i) % (1 << 31)
    return (-(init + step * (i - half))) % (1 << 31)


def bitrev(i, n):
    return int(format(i, "0%db" % n)[::-1], 2) if n else i


def double_x(x):
    return (2 * x * x - 1) % P


def circle_basis_eval(coeffs, x, y):
    """sum_i c_i * prod_{bit j of i} f_j with f = [y, x, pi(x), pi^2(x), ...] (M31 point)."""
    m = len(coeffs).bit_length() - 1
    f = [y]
    if m > 1:
        f.append(x)
        for _ in range(2, m):
            f.append(double_x(f[-1]))
    total = 0
    for i, c in enumerate(coeffs):
        t = c
        for j in range(m):
            if (i >> j) & 1:
                t = t * f[j] % P
        total += t
    return total % P
