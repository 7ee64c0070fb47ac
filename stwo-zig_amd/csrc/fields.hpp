// Exact M31 / CM31 / QM31 arithmetic and the M31 circle group, usable from
// host and device code.
//
// Field laws follow src/core/fields/m31.zig:49-171 (canonical [0,p) words,
// one-fold product reduction), cm31.zig:97-129 and qm31.zig:13,153-222,269-284
// (u^2 = 2+i, mulByR = (2a-b, a+2b)).  The reference multiplies CM31 by Karatsuba;
// here the CM31 and QM31 products are written out as raw products reduced once per
// coordinate (the same field elements).  The circle group follows
// src/core/circle.zig:8-374.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SR_HD __host__ __device__ __forceinline__

// Global-address-space views of device pointers read from pointer tables, so
// that loads and stores compile to global_* rather than flat_* instructions
// (flat accesses also count against the LDS wait counter).
#define SR_GLOBAL __attribute__((address_space(1)))
typedef SR_GLOBAL uint32_t gu32;
typedef SR_GLOBAL const uint32_t gcu32;
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));  // 16-byte global accesses

namespace sr {

constexpr uint32_t P = 0x7fffffffu;

// Canonical add/sub in three ops: the unsigned minimum picks the reduced
// candidate because the other one wraps above 2^31.
SR_HD uint32_t umin32(uint32_t a, uint32_t b) { return a < b ? a : b; }
SR_HD uint32_t madd(uint32_t a, uint32_t b) {
    uint32_t s = a + b;
    return umin32(s, s - P);
}
SR_HD uint32_t msub(uint32_t a, uint32_t b) {
    uint32_t d = a - b;
    return umin32(d, d + P);
}
SR_HD uint32_t mneg(uint32_t a) { return a == 0 ? 0 : P - a; }
// Product of two canonical words is < p^2 < 2^62: one Mersenne fold lands in
// [0, 2p) and a conditional subtraction canonicalises (m31.zig:164-171).
SR_HD uint32_t mmul(uint32_t a, uint32_t b) {
    uint64_t x = (uint64_t)a * b;
    uint32_t r = (uint32_t)(x & P) + (uint32_t)(x >> 31);
    return umin32(r, r - P);
}
// b * t mod p for canonical b, t given t2 = 2t (a plain u32): b*t2 = 2bt splits as
// hi = floor(bt / 2^31) and lo / 2 = bt mod 2^31, so bt = hi + (lo >> 1) (mod p) with
// hi + (lo >> 1) < 2p; one conditional subtraction canonicalises.
SR_HD uint32_t mmul_t2(uint32_t b, uint32_t t2) {
    uint64_t x = (uint64_t)b * t2;
    uint32_t r = (uint32_t)(x >> 32) + ((uint32_t)x >> 1);
    return umin32(r, r - P);
}
// 2 * (p - w) for a doubled word w2 = 2w: the doubled negation (0 stays 0).
SR_HD uint32_t mneg_t2(uint32_t w2) { return w2 == 0 ? 0u : 0xFFFFFFFEu - w2; }
// Fold a 64-bit accumulator of folded products to a canonical word (m31.zig:150-158).
SR_HD uint32_t mred64(uint64_t x) {
    uint64_t t = (x & P) + (x >> 31);
    t = (t & P) + (t >> 31);
    uint32_t r = (uint32_t)t;
    return r >= P ? r - P : r;
}
// One Mersenne fold of a product, < 2^32; 2^32 such terms fit a u64 accumulator.
SR_HD uint64_t mfold(uint64_t x) { return (x & P) + (x >> 31); }
SR_HD uint32_t msqr(uint32_t a) { return mmul(a, a); }
SR_HD uint32_t mpow(uint32_t a, uint32_t e) {
    uint32_t acc = 1;
    while (e) {
        if (e & 1) acc = mmul(acc, a);
        a = mmul(a, a);
        e >>= 1;
    }
    return acc;
}
// a^(p-2); p-2 = 2^31 - 3.  Inverse is unique, so any exponentiation chain is exact.
// Addition chain over the exponents 2^k - 1 (x_k = a^(2^k - 1), x_{j+k} = x_j^(2^k) x_k):
// x2, x4, x8, x16, x24, x28, x29, then a^(2^31 - 3) = x29^4 a.  30 squarings and 8
// multiplications.
SR_HD uint32_t msqrn(uint32_t a, int n) {
    for (int i = 0; i < n; i++) a = msqr(a);
    return a;
}
SR_HD uint32_t minv(uint32_t a) {
    const uint32_t x2 = mmul(msqr(a), a);
    const uint32_t x4 = mmul(msqrn(x2, 2), x2);
    const uint32_t x8 = mmul(msqrn(x4, 4), x4);
    const uint32_t x16 = mmul(msqrn(x8, 8), x8);
    const uint32_t x24 = mmul(msqrn(x16, 8), x8);
    const uint32_t x28 = mmul(msqrn(x24, 4), x4);
    const uint32_t x29 = mmul(msqr(x28), a);
    return mmul(msqrn(x29, 2), a);
}

struct cm31 {
    uint32_t a, b;
};
SR_HD cm31 cadd(cm31 x, cm31 y) { return {madd(x.a, y.a), madd(x.b, y.b)}; }
SR_HD cm31 csub(cm31 x, cm31 y) { return {msub(x.a, y.a), msub(x.b, y.b)}; }
SR_HD cm31 cneg(cm31 x) { return {mneg(x.a), mneg(x.b)}; }
// (a + bi)(c + di) = (ac - bd) + (ad + bc)i: each coordinate is a sum of two raw 62-bit
// products (the subtraction as b (p - d)) reduced once.
SR_HD cm31 cmul(cm31 x, cm31 y) {
    const uint64_t re = (uint64_t)x.a * y.a + (uint64_t)x.b * (P - y.b);
    const uint64_t im = (uint64_t)x.a * y.b + (uint64_t)x.b * y.a;
    return {mred64(re), mred64(im)};
}
SR_HD cm31 cmulm(cm31 x, uint32_t m) { return {mmul(x.a, m), mmul(x.b, m)}; }
SR_HD cm31 cinv(cm31 x) {
    uint32_t d = minv(madd(mmul(x.a, x.a), mmul(x.b, x.b)));
    return {mmul(x.a, d), mmul(mneg(x.b), d)};
}
SR_HD cm31 cmul_r(cm31 v) { return {msub(madd(v.a, v.a), v.b), madd(v.a, madd(v.b, v.b))}; }

struct qm31 {
    cm31 c0, c1;
};
SR_HD qm31 qzero() { return {{0, 0}, {0, 0}}; }
SR_HD qm31 qone() { return {{1, 0}, {0, 0}}; }
SR_HD qm31 qfromm(uint32_t m) { return {{m, 0}, {0, 0}}; }
SR_HD qm31 qadd(qm31 x, qm31 y) { return {cadd(x.c0, y.c0), cadd(x.c1, y.c1)}; }
SR_HD qm31 qsub(qm31 x, qm31 y) { return {csub(x.c0, y.c0), csub(x.c1, y.c1)}; }
SR_HD qm31 qneg(qm31 x) { return {cneg(x.c0), cneg(x.c1)}; }
// QM31 = CM31[u] / (u^2 - (2 + i)).  With x = (a0 + b0 i) + (a1 + b1 i) u and
// y = (c0 + d0 i) + (c1 + d1 i) u, x y = x0 y0 + (2 + i) x1 y1 + (x0 y1 + x1 y0) u, whose four
// coordinates, grouped by the words of x, are
//   r0 = a0 c0 - b0 d0 + a1 (2 c1 - d1) - b1 (c1 + 2 d1)
//   r1 = a0 d0 + b0 c0 + a1 (c1 + 2 d1) + b1 (2 c1 - d1)
//   r2 = a0 c1 - b0 d1 + a1 c0 - b1 d0
//   r3 = a0 d1 + b0 c1 + a1 d0 + b1 c0.
// Negated factors are taken as p - v (at most p), so each coordinate is four raw products
// (< 2^62 each, < 2^64 together) reduced once: 16 multiply-adds and 4 reductions instead of
// 9 reduced products and 15 reduced additions of the Karatsuba form.  The combinations of
// y's words (t, s, the negations) come first, so a uniform or loop-invariant factor goes in y.
SR_HD qm31 qmul(qm31 x, qm31 y) {
    const uint32_t a0 = x.c0.a, b0 = x.c0.b, a1 = x.c1.a, b1 = x.c1.b;
    const uint32_t c0 = y.c0.a, d0 = y.c0.b, c1 = y.c1.a, d1 = y.c1.b;
    const uint32_t nd0 = P - d0, nd1 = P - d1;
    const uint32_t t = madd(madd(c1, c1), nd1);  // 2 c1 - d1
    const uint32_t s = madd(c1, madd(d1, d1));   // c1 + 2 d1
    const uint64_t r0 = (uint64_t)a0 * c0 + (uint64_t)b0 * nd0 + (uint64_t)a1 * t + (uint64_t)b1 * (P - s);
    const uint64_t r1 = (uint64_t)a0 * d0 + (uint64_t)b0 * c0 + (uint64_t)a1 * s + (uint64_t)b1 * t;
    const uint64_t r2 = (uint64_t)a0 * c1 + (uint64_t)b0 * nd1 + (uint64_t)a1 * c0 + (uint64_t)b1 * nd0;
    const uint64_t r3 = (uint64_t)a0 * d1 + (uint64_t)b0 * c1 + (uint64_t)a1 * d0 + (uint64_t)b1 * c0;
    return {{mred64(r0), mred64(r1)}, {mred64(r2), mred64(r3)}};
}
SR_HD qm31 qmulm(qm31 x, uint32_t m) { return {cmulm(x.c0, m), cmulm(x.c1, m)}; }
SR_HD qm31 qmulc(qm31 x, cm31 c) { return {cmul(x.c0, c), cmul(x.c1, c)}; }
SR_HD qm31 qsqr(qm31 x) { return qmul(x, x); }
SR_HD qm31 qconj(qm31 x) { return {x.c0, cneg(x.c1)}; }
SR_HD bool qeq(qm31 x, qm31 y) {
    return x.c0.a == y.c0.a && x.c0.b == y.c0.b && x.c1.a == y.c1.a && x.c1.b == y.c1.b;
}
SR_HD bool qiszero(qm31 x) { return (x.c0.a | x.c0.b | x.c1.a | x.c1.b) == 0; }
SR_HD qm31 qinv(qm31 x) {
    cm31 b2 = cmul(x.c1, x.c1);
    cm31 rb2 = cadd(cadd(b2, b2), cm31{mneg(b2.b), b2.a});
    cm31 den = csub(cmul(x.c0, x.c0), rb2);
    cm31 di = cinv(den);
    return {cmul(x.c0, di), cmul(cneg(x.c1), di)};
}
SR_HD qm31 qload(const uint32_t* w) { return {{w[0], w[1]}, {w[2], w[3]}}; }
SR_HD void qstore(qm31 x, uint32_t* w) {
    w[0] = x.c0.a;
    w[1] = x.c0.b;
    w[2] = x.c1.a;
    w[3] = x.c1.b;
}

// ---- circle group over M31 (circle.zig) ----
struct pt {
    uint32_t x, y;
};
SR_HD pt padd(pt a, pt b) {
    return {msub(mmul(a.x, b.x), mmul(a.y, b.y)), madd(mmul(a.x, b.y), mmul(a.y, b.x))};
}
SR_HD pt pconj(pt a) { return {a.x, mneg(a.y)}; }
SR_HD uint32_t mdouble_x(uint32_t x) {
    uint32_t s = msqr(x);
    return msub(madd(s, s), 1);
}
constexpr uint32_t ORDER_MASK = 0x7fffffffu;
SR_HD uint32_t idx_neg(uint32_t i) { return (0x80000000u - (i & ORDER_MASK)) & ORDER_MASK; }
SR_HD uint32_t subgroup_gen(uint32_t log_size) { return 1u << (31 - log_size); }

// Generator doubles G_k = 2^k * (2, 1268011823) (circle.zig:137-143,226-237).
struct GenDoubles {
    pt d[31];
};
SR_HD GenDoubles make_gen_doubles() {
    GenDoubles g;
    pt cur = {2u, 1268011823u};
    for (int i = 0; i < 31; i++) {
        g.d[i] = cur;
        cur = padd(cur, cur);
    }
    return g;
}
// CirclePointIndex.toPoint (circle.zig:187-196).
SR_HD pt idx_to_point(const GenDoubles& g, uint32_t idx) {
    pt acc = {1, 0};
    uint32_t bits = idx & ORDER_MASK;
    while (bits) {
        int b = __builtin_ctz(bits);
        acc = padd(acc, g.d[b]);
        bits &= bits - 1;
    }
    return acc;
}
// Coset.halfOdds(log) = (subgroupGen(log+2), step subgroupGen(log)) (circle.zig:280-282).
SR_HD uint32_t half_odds_initial(uint32_t log_size) { return subgroup_gen(log_size + 2); }
// CanonicCoset(log).circleDomain().indexAt(i) (canonic.zig:24-26; domain.zig:58-62).
SR_HD uint32_t domain_index_at(uint32_t log_size, uint64_t i) {
    uint32_t init = half_odds_initial(log_size - 1), step = subgroup_gen(log_size - 1);
    uint64_t half = 1ull << (log_size - 1);
    if (i < half) return (uint32_t)((init + (uint64_t)step * i) & ORDER_MASK);
    return idx_neg((uint32_t)((init + (uint64_t)step * (i - half)) & ORDER_MASK));
}
SR_HD uint32_t bitrev32(uint32_t i, uint32_t log_size) {
    if (log_size == 0) return i;
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_bitreverse32(i) >> (32 - log_size);
#else
    uint32_t r = 0;
    for (uint32_t k = 0; k < log_size; k++) r |= ((i >> k) & 1u) << (log_size - 1 - k);
    return r;
#endif
}

}  // namespace sr
