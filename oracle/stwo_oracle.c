/*
 * stwo_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Plain-C restatement of the reference CPU backend's wide-Fibonacci proving
 * path (welttowelt/stwo-zig, pinned upstream Stwo a8fcf4bd).  Loaded only by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   - every function below is exercised by the end-to-end byte equality of the
 *     six committed golden proofs (5x16 pow0, 10x8, 14x32, 16x64, 18x100, 20x100);
 *   - fields: closed forms / Python big-int brute force;
 *   - Blake2s: hashlib.blake2s and the four Merkle KATs of
 *     src/core/vcs_lifted/blake2_merkle.zig:340-384;
 *   - FFT: interpolate(evaluate(x)) == x and equality with direct O(n^2)
 *     evaluation of the circle basis at domain points;
 *   - FRI folds: a folded low-degree evaluation stays low degree.
 * Threading: OpenMP over columns / rows; results are order independent.
 */
#include "stwo_oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define P ORA_P

/* ======================================================================
 * M31 (src/core/fields/m31.zig:10-171)
 * ==================================================================== */
static inline uint32_t m_add(uint32_t a, uint32_t b) {
    uint32_t s = a + b;
    return s >= P ? s - P : s;
}
static inline uint32_t m_sub(uint32_t a, uint32_t b) { return a >= b ? a - b : a + P - b; }
static inline uint32_t m_neg(uint32_t a) { return a == 0 ? 0 : P - a; }
/* reduceProduct (m31.zig:164-171): one Mersenne fold + conditional subtract. */
static inline uint32_t m_mul(uint32_t a, uint32_t b) {
    uint64_t x = (uint64_t)a * b;
    uint32_t r = (uint32_t)((x & P) + (x >> 31));
    return r >= P ? r - P : r;
}
/* reduce64 (m31.zig:150-158). */
static inline uint32_t m_from_u64(uint64_t x) {
    uint64_t t = (x & P) + (x >> 31);
    t = (t & P) + (t >> 31);
    uint32_t r = (uint32_t)t;
    uint32_t r2 = r - P;
    return r < r2 ? r : r2;
}
static uint32_t m_pow(uint32_t a, uint64_t e) {
    uint32_t acc = 1;
    while (e) {
        if (e & 1) acc = m_mul(acc, a);
        a = m_mul(a, a);
        e >>= 1;
    }
    return acc;
}
/* invUncheckedNonZero = a^(p-2) (m31.zig:97-100). */
static inline uint32_t m_inv(uint32_t a) { return m_pow(a, P - 2); }

uint32_t ora_m31_mul(uint32_t a, uint32_t b) { return m_mul(a, b); }
uint32_t ora_m31_inv(uint32_t a) { return m_inv(a); }

/* ======================================================================
 * CM31 = M31[i]/(i^2+1) (src/core/fields/cm31.zig:97-129)
 * ==================================================================== */
typedef struct { uint32_t a, b; } cm31;
static inline cm31 c_add(cm31 x, cm31 y) { return (cm31){m_add(x.a, y.a), m_add(x.b, y.b)}; }
static inline cm31 c_sub(cm31 x, cm31 y) { return (cm31){m_sub(x.a, y.a), m_sub(x.b, y.b)}; }
static inline cm31 c_neg(cm31 x) { return (cm31){m_neg(x.a), m_neg(x.b)}; }
static inline cm31 c_mul(cm31 x, cm31 y) {
    uint32_t ac = m_mul(x.a, y.a), bd = m_mul(x.b, y.b);
    uint32_t cross = m_mul(m_add(x.a, x.b), m_add(y.a, y.b));
    return (cm31){m_sub(ac, bd), m_sub(m_sub(cross, ac), bd)};
}
static inline cm31 c_mul_m(cm31 x, uint32_t m) { return (cm31){m_mul(x.a, m), m_mul(x.b, m)}; }
static inline cm31 c_square(cm31 x) { return c_mul(x, x); }
static inline int c_is_zero(cm31 x) { return x.a == 0 && x.b == 0; }
/* (a+bi)^-1 = (a-bi)/(a^2+b^2) (cm31.zig inv). */
static inline cm31 c_inv(cm31 x) {
    uint32_t d = m_inv(m_add(m_mul(x.a, x.a), m_mul(x.b, x.b)));
    return (cm31){m_mul(x.a, d), m_mul(m_neg(x.b), d)};
}
/* mulByR: (a+bi)(2+i) = (2a-b) + (a+2b)i (qm31.zig:269-277). */
static inline cm31 c_mul_r(cm31 v) {
    return (cm31){m_sub(m_add(v.a, v.a), v.b), m_add(v.a, m_add(v.b, v.b))};
}

/* ======================================================================
 * QM31 = CM31[u]/(u^2 - (2+i)) (src/core/fields/qm31.zig:13-222)
 * words (c0.a, c0.b, c1.a, c1.b) (qm31.zig:64-70)
 * ==================================================================== */
typedef struct { cm31 c0, c1; } qm31;
static const qm31 Q_ZERO = {{0, 0}, {0, 0}};
static const qm31 Q_ONE = {{1, 0}, {0, 0}};
static inline qm31 q_from_m(uint32_t m) { return (qm31){{m, 0}, {0, 0}}; }
static inline qm31 q_add(qm31 x, qm31 y) { return (qm31){c_add(x.c0, y.c0), c_add(x.c1, y.c1)}; }
static inline qm31 q_sub(qm31 x, qm31 y) { return (qm31){c_sub(x.c0, y.c0), c_sub(x.c1, y.c1)}; }
static inline qm31 q_neg(qm31 x) { return (qm31){c_neg(x.c0), c_neg(x.c1)}; }
static inline qm31 q_mul(qm31 x, qm31 y) {
    cm31 ac = c_mul(x.c0, y.c0), bd = c_mul(x.c1, y.c1);
    cm31 cross = c_sub(c_sub(c_mul(c_add(x.c0, x.c1), c_add(y.c0, y.c1)), ac), bd);
    return (qm31){c_add(ac, c_mul_r(bd)), cross};
}
static inline qm31 q_mul_m(qm31 x, uint32_t m) { return (qm31){c_mul_m(x.c0, m), c_mul_m(x.c1, m)}; }
static inline qm31 q_mul_c(qm31 x, cm31 c) { return (qm31){c_mul(x.c0, c), c_mul(x.c1, c)}; }
static inline qm31 q_square(qm31 x) { return q_mul(x, x); }
static inline int q_eq(qm31 x, qm31 y) {
    return x.c0.a == y.c0.a && x.c0.b == y.c0.b && x.c1.a == y.c1.a && x.c1.b == y.c1.b;
}
static inline int q_is_zero(qm31 x) { return c_is_zero(x.c0) && c_is_zero(x.c1); }
/* (a+bu)^-1 = (a-bu)/(a^2-(2+i)b^2) (qm31.zig:182-199). */
static qm31 q_inv(qm31 x) {
    cm31 b2 = c_square(x.c1);
    cm31 rb2 = c_add(c_add(b2, b2), (cm31){m_neg(b2.b), b2.a});
    cm31 denom = c_sub(c_square(x.c0), rb2);
    cm31 di = c_inv(denom);
    return (qm31){c_mul(x.c0, di), c_mul(c_neg(x.c1), di)};
}
/* complexConjugate: (c0, -c1) (qm31.zig:205-210). */
static inline qm31 q_conj(qm31 x) { return (qm31){x.c0, c_neg(x.c1)}; }
static inline void q_store(qm31 x, uint32_t *w) {
    w[0] = x.c0.a; w[1] = x.c0.b; w[2] = x.c1.a; w[3] = x.c1.b;
}
static inline qm31 q_load(const uint32_t *w) { return (qm31){{w[0], w[1]}, {w[2], w[3]}}; }

void ora_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]) {
    q_store(q_mul(q_load(a), q_load(b)), out);
}
void ora_qm31_inv(const uint32_t a[4], uint32_t out[4]) { q_store(q_inv(q_load(a)), out); }

/* ======================================================================
 * Circle group (src/core/circle.zig:8-374)
 * ==================================================================== */
typedef struct { uint32_t x, y; } pt;
typedef struct { qm31 x, y; } qpt;
#define LOG_ORDER 31
#define ORDER_MASK 0x7fffffffu

static inline pt p_add(pt a, pt b) {
    return (pt){m_sub(m_mul(a.x, b.x), m_mul(a.y, b.y)), m_add(m_mul(a.x, b.y), m_mul(a.y, b.x))};
}
static inline pt p_conj(pt a) { return (pt){a.x, m_neg(a.y)}; }
static inline qpt qp_add(qpt a, qpt b) {
    return (qpt){q_sub(q_mul(a.x, b.x), q_mul(a.y, b.y)), q_add(q_mul(a.x, b.y), q_mul(a.y, b.x))};
}
static inline qm31 q_double_x(qm31 x) {
    qm31 s = q_square(x);
    return q_sub(q_add(s, s), Q_ONE);
}
static inline uint32_t m_double_x(uint32_t x) {
    uint32_t s = m_mul(x, x);
    return m_sub(m_add(s, s), 1);
}

static pt GEN_DOUBLES[LOG_ORDER];
static int g_init_done = 0;
static void init_gen_doubles(void) {
    if (g_init_done) return;
    /* M31_CIRCLE_GEN = (2, 1268011823), log order 31 (circle.zig:137-143). */
    pt cur = {2, 1268011823u};
    for (int i = 0; i < LOG_ORDER; i++) {
        GEN_DOUBLES[i] = cur;
        cur = p_add(cur, cur);
    }
    g_init_done = 1;
}
/* CirclePointIndex.toPoint (circle.zig:187-196). */
static pt idx_to_point(uint32_t idx) {
    init_gen_doubles();
    pt acc = {1, 0};
    uint32_t bits = idx & ORDER_MASK;
    while (bits) {
        int b = __builtin_ctz(bits);
        acc = p_add(acc, GEN_DOUBLES[b]);
        bits &= bits - 1;
    }
    return acc;
}
static inline uint32_t idx_neg(uint32_t i) { return (0x80000000u - (i & ORDER_MASK)) & ORDER_MASK; }
static inline uint32_t subgroup_gen(uint32_t log_size) { return 1u << (LOG_ORDER - log_size); }

/* Coset (circle.zig:245-374). */
typedef struct { uint32_t initial_index, step_size, log_size; } coset;
static inline coset coset_new(uint32_t initial, uint32_t log_size) {
    return (coset){initial & ORDER_MASK, subgroup_gen(log_size), log_size};
}
static inline coset coset_half_odds(uint32_t log_size) { return coset_new(subgroup_gen(log_size + 2), log_size); }
static inline coset coset_odds(uint32_t log_size) { return coset_new(subgroup_gen(log_size + 1), log_size); }
static inline coset coset_double(coset c) {
    return (coset){(c.initial_index * 2u) & ORDER_MASK, (c.step_size * 2u) & ORDER_MASK, c.log_size - 1};
}
static inline uint32_t coset_index_at(coset c, uint64_t i) {
    return (uint32_t)((c.initial_index + (uint64_t)c.step_size * i) & ORDER_MASK);
}
static inline pt coset_at(coset c, uint64_t i) { return idx_to_point(coset_index_at(c, i)); }

/* CanonicCoset(log).circleDomain() = CircleDomain(halfOdds(log-1)) (canonic.zig:9-28);
 * CircleDomain.indexAt (domain.zig:58-62). */
static inline uint32_t domain_index_at(uint32_t log_size, uint64_t i) {
    coset hc = coset_half_odds(log_size - 1);
    uint64_t half = 1ull << (log_size - 1);
    if (i < half) return coset_index_at(hc, i);
    return idx_neg(coset_index_at(hc, i - half));
}
static inline pt domain_at(uint32_t log_size, uint64_t i) { return idx_to_point(domain_index_at(log_size, i)); }

void ora_circle_point_at_index(uint32_t idx, uint32_t out[2]) {
    pt p = idx_to_point(idx);
    out[0] = p.x; out[1] = p.y;
}
void ora_circle_domain_point(uint32_t log_size, uint64_t i, uint32_t out[2]) {
    pt p = domain_at(log_size, i);
    out[0] = p.x; out[1] = p.y;
}

/* bitReverseIndex (core/utils.zig:14-18). */
static inline uint64_t bitrev(uint64_t i, uint32_t log_size) {
    if (log_size == 0) return i;
    uint64_t r = 0;
    for (uint32_t k = 0; k < log_size; k++) r |= ((i >> k) & 1ull) << (log_size - 1 - k);
    return r;
}
static void bitrev_q(qm31 *v, uint64_t n) {
    uint32_t lg = 0;
    while ((1ull << lg) < n) lg++;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t j = bitrev(i, lg);
        if (j > i) { qm31 t = v[i]; v[i] = v[j]; v[j] = t; }
    }
}

/* ======================================================================
 * Twiddles (src/prover/poly/twiddles.zig:37-91)
 * ==================================================================== */
void ora_twiddles(uint32_t circle_log, uint32_t *tw, uint32_t *itw) {
    coset c = coset_half_odds(circle_log - 1); /* CanonicCoset(log).circleDomain().half_coset */
    uint64_t at = 0;
    for (uint32_t i = 0; i < circle_log - 1; i++) {
        uint64_t layer_len = (1ull << c.log_size) / 2;
        uint32_t lg = c.log_size - 1;
        pt cur = idx_to_point(c.initial_index), step = idx_to_point(c.step_size);
        for (uint64_t j = 0; j < layer_len; j++) {
            tw[at + bitrev(j, lg)] = cur.x;
            cur = p_add(cur, step);
        }
        at += layer_len;
        c = coset_double(c);
    }
    tw[at] = 1;
    if (itw)
        for (uint64_t j = 0; j <= at; j++) itw[j] = m_inv(tw[j]);
}

/* ======================================================================
 * Circle FFT (src/prover/poly/circle/transforms.zig:149-360; core/fft.zig:8-18)
 * Values are stored in bit-reversed circle-domain order.
 * ==================================================================== */
static inline void bfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t tmp = m_mul(*v1, t);
    *v1 = m_sub(*v0, tmp);
    *v0 = m_add(*v0, tmp);
}
static inline void ibfly(uint32_t *v0, uint32_t *v1, uint32_t t) {
    uint32_t tmp = *v0;
    *v0 = m_add(tmp, *v1);
    *v1 = m_mul(m_sub(tmp, *v1), t);
}

typedef struct { uint32_t log; uint32_t *tw, *itw; } twiddle_tree;
static twiddle_tree tw_cache[32];
static twiddle_tree *get_twiddles(uint32_t log) {
    twiddle_tree *t = &tw_cache[log];
    if (!t->tw) {
        uint64_t n = 1ull << (log - 1);
        uint32_t *tw = malloc(n * 4), *itw = malloc(n * 4);
        ora_twiddles(log, tw, itw);
        t->log = log;
        t->itw = itw;
        __atomic_store_n(&t->tw, tw, __ATOMIC_RELEASE);
    }
    return t;
}
/* Layer j (stage j+1) slice: twiddles[len_total - 2^(m-1-j) .. ] of length 2^(m-2-j). */
static inline const uint32_t *tw_layer(const uint32_t *tw, uint32_t m, uint32_t j) {
    return tw + ((1ull << (m - 1)) - (1ull << (m - 1 - j)));
}

/* evaluateBufferWithTwiddles (transforms.zig:149-246). */
static void circle_evaluate(uint32_t *v, uint32_t m) {
    if (m == 1) {
        pt p0 = idx_to_point(coset_half_odds(0).initial_index);
        bfly(&v[0], &v[1], p0.y);
        return;
    }
    if (m == 2) {
        pt p0 = idx_to_point(coset_half_odds(1).initial_index);
        bfly(&v[0], &v[2], p0.x);
        bfly(&v[1], &v[3], p0.x);
        bfly(&v[0], &v[1], p0.y);
        bfly(&v[2], &v[3], m_neg(p0.y));
        return;
    }
    const uint32_t *tw = get_twiddles(m)->tw;
    for (uint32_t s = m - 1; s >= 1; s--) {
        const uint32_t *T = tw_layer(tw, m, s - 1);
        uint64_t nblocks = 1ull << (m - 1 - s), half = 1ull << s;
        for (uint64_t h = 0; h < nblocks; h++) {
            uint32_t t = T[h];
            uint32_t *b = v + (h << (s + 1));
            for (uint64_t l = 0; l < half; l++) bfly(&b[l], &b[l + half], t);
        }
    }
    const uint32_t *T0 = tw_layer(tw, m, 0);
    for (uint64_t k = 0; k < (1ull << (m - 3)); k++) {
        uint32_t x = T0[2 * k], y = T0[2 * k + 1];
        uint32_t *b = v + 8 * k;
        bfly(&b[0], &b[1], y);
        bfly(&b[2], &b[3], m_neg(y));
        bfly(&b[4], &b[5], m_neg(x));
        bfly(&b[6], &b[7], x);
    }
}

/* interpolateIntoBufferWithTwiddles (transforms.zig:249-360). */
static void circle_interpolate(uint32_t *v, uint32_t m) {
    if (m == 1) {
        pt p0 = idx_to_point(coset_half_odds(0).initial_index);
        uint32_t y = p0.y;
        uint32_t yn_inv = m_inv(m_mul(y, 2));
        uint32_t y_inv = m_mul(yn_inv, 2), n_inv = m_mul(yn_inv, y);
        ibfly(&v[0], &v[1], y_inv);
        v[0] = m_mul(v[0], n_inv);
        v[1] = m_mul(v[1], n_inv);
        return;
    }
    if (m == 2) {
        pt p0 = idx_to_point(coset_half_odds(1).initial_index);
        uint32_t x = p0.x, y = p0.y;
        uint32_t xyn_inv = m_inv(m_mul(m_mul(x, y), 4));
        uint32_t x_inv = m_mul(m_mul(xyn_inv, y), 4), y_inv = m_mul(m_mul(xyn_inv, x), 4);
        uint32_t n_inv = m_mul(m_mul(xyn_inv, x), y);
        ibfly(&v[0], &v[1], y_inv);
        ibfly(&v[2], &v[3], m_neg(y_inv));
        ibfly(&v[0], &v[2], x_inv);
        ibfly(&v[1], &v[3], x_inv);
        for (int i = 0; i < 4; i++) v[i] = m_mul(v[i], n_inv);
        return;
    }
    const uint32_t *itw = get_twiddles(m)->itw;
    const uint32_t *T0 = tw_layer(itw, m, 0);
    for (uint64_t k = 0; k < (1ull << (m - 3)); k++) {
        uint32_t x = T0[2 * k], y = T0[2 * k + 1];
        uint32_t *b = v + 8 * k;
        ibfly(&b[0], &b[1], y);
        ibfly(&b[2], &b[3], m_neg(y));
        ibfly(&b[4], &b[5], m_neg(x));
        ibfly(&b[6], &b[7], x);
    }
    for (uint32_t s = 1; s < m; s++) {
        const uint32_t *T = tw_layer(itw, m, s - 1);
        uint64_t nblocks = 1ull << (m - 1 - s), half = 1ull << s;
        for (uint64_t h = 0; h < nblocks; h++) {
            uint32_t t = T[h];
            uint32_t *b = v + (h << (s + 1));
            for (uint64_t l = 0; l < half; l++) ibfly(&b[l], &b[l + half], t);
        }
    }
    uint32_t n_inv = m_inv((uint32_t)((1ull << m) % P));
    for (uint64_t i = 0; i < (1ull << m); i++) v[i] = m_mul(v[i], n_inv);
}

void ora_circle_evaluate(uint32_t *values, uint32_t log_size) { circle_evaluate(values, log_size); }
void ora_circle_interpolate(uint32_t *values, uint32_t log_size) { circle_interpolate(values, log_size); }
/* evaluateExtensionBufferWithTwiddles (transforms.zig:132-147): zero padding. */
void ora_circle_lde(const uint32_t *coeffs, uint32_t log_size, uint32_t log_blowup, uint32_t *out) {
    uint64_t n = 1ull << log_size, N = n << log_blowup;
    memcpy(out, coeffs, n * 4);
    memset(out + n, 0, (N - n) * 4);
    circle_evaluate(out, log_size + log_blowup);
}

/* ======================================================================
 * Point evaluation (point_evaluation.zig:14-38,469-488)
 * p(P) = sum_i c_i * prod_{bit j of i} f_j, f = [P.y, P.x, 2x^2-1, ...].
 * Evaluated through the weight vector w_i = prod f_j^{bit_j(i)}.
 * ==================================================================== */
static qm31 *eval_weights(qpt point, uint32_t m) {
    qm31 f[32];
    f[0] = point.y;
    if (m > 1) {
        f[1] = point.x;
        for (uint32_t k = 2; k < m; k++) f[k] = q_double_x(f[k - 1]);
    }
    uint64_t n = 1ull << m;
    qm31 *w = malloc(n * sizeof(qm31));
    w[0] = Q_ONE;
    for (uint32_t k = 0; k < m; k++) {
        uint64_t half = 1ull << k;
        for (uint64_t i = 0; i < half; i++) w[half + i] = q_mul(w[i], f[k]);
    }
    return w;
}
static qm31 dot_weights(const uint32_t *c, const qm31 *w, uint64_t n) {
    uint64_t acc[4] = {0, 0, 0, 0};
    qm31 r = Q_ZERO;
    for (uint64_t i = 0; i < n; i++) {
        acc[0] += (uint64_t)c[i] * w[i].c0.a;
        acc[1] += (uint64_t)c[i] * w[i].c0.b;
        acc[2] += (uint64_t)c[i] * w[i].c1.a;
        acc[3] += (uint64_t)c[i] * w[i].c1.b;
        if ((i & 3) == 3) {
            r = q_add(r, (qm31){{m_from_u64(acc[0]), m_from_u64(acc[1])}, {m_from_u64(acc[2]), m_from_u64(acc[3])}});
            acc[0] = acc[1] = acc[2] = acc[3] = 0;
        }
    }
    r = q_add(r, (qm31){{m_from_u64(acc[0]), m_from_u64(acc[1])}, {m_from_u64(acc[2]), m_from_u64(acc[3])}});
    return r;
}
void ora_eval_at_point(const uint32_t *coeffs, uint32_t log_size, const uint32_t point[8], uint32_t out[4]) {
    qpt P_ = {q_load(point), q_load(point + 4)};
    if (log_size == 0) {
        q_store(q_from_m(coeffs[0]), out);
        return;
    }
    qm31 *w = eval_weights(P_, log_size);
    q_store(dot_weights(coeffs, w, 1ull << log_size), out);
    free(w);
}

/* ======================================================================
 * Blake2s-256, unkeyed (RFC 7693; src/core/crypto/blake2s_backend.zig)
 * ==================================================================== */
static const uint32_t B2_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                  0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t B2_SIGMA[10][16] = {
    {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15}, {14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3},
    {11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4}, {7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8},
    {9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13}, {2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9},
    {12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11}, {13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10},
    {6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5}, {10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0}};
static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
#define B2G(a, b, c, d, x, y)            \
    do {                                 \
        a = a + b + x; d = rotr(d ^ a, 16); \
        c = c + d; b = rotr(b ^ c, 12);  \
        a = a + b + y; d = rotr(d ^ a, 8);  \
        c = c + d; b = rotr(b ^ c, 7);   \
    } while (0)
static void b2_compress(uint32_t h[8], const uint32_t m[16], uint64_t t, int last) {
    uint32_t v[16];
    for (int i = 0; i < 8; i++) { v[i] = h[i]; v[i + 8] = B2_IV[i]; }
    v[12] ^= (uint32_t)t;
    v[13] ^= (uint32_t)(t >> 32);
    if (last) v[14] = ~v[14];
    for (int r = 0; r < 10; r++) {
        const uint8_t *s = B2_SIGMA[r];
        B2G(v[0], v[4], v[8], v[12], m[s[0]], m[s[1]]);
        B2G(v[1], v[5], v[9], v[13], m[s[2]], m[s[3]]);
        B2G(v[2], v[6], v[10], v[14], m[s[4]], m[s[5]]);
        B2G(v[3], v[7], v[11], v[15], m[s[6]], m[s[7]]);
        B2G(v[0], v[5], v[10], v[15], m[s[8]], m[s[9]]);
        B2G(v[1], v[6], v[11], v[12], m[s[10]], m[s[11]]);
        B2G(v[2], v[7], v[8], v[13], m[s[12]], m[s[13]]);
        B2G(v[3], v[4], v[9], v[14], m[s[14]], m[s[15]]);
    }
    for (int i = 0; i < 8; i++) h[i] ^= v[i] ^ v[i + 8];
}
typedef struct { uint32_t h[8]; uint8_t buf[64]; uint32_t buflen; uint64_t t; } b2s;
static void b2_init(b2s *s) {
    memcpy(s->h, B2_IV, 32);
    s->h[0] ^= 0x01010020u; /* digest 32, no key, fanout 1, depth 1 */
    s->buflen = 0;
    s->t = 0;
}
static void b2_update(b2s *s, const uint8_t *d, size_t n) {
    while (n > 0) {
        if (s->buflen == 64) {
            s->t += 64;
            uint32_t m[16];
            memcpy(m, s->buf, 64);
            b2_compress(s->h, m, s->t, 0);
            s->buflen = 0;
        }
        size_t take = 64 - s->buflen;
        if (take > n) take = n;
        memcpy(s->buf + s->buflen, d, take);
        s->buflen += (uint32_t)take;
        d += take;
        n -= take;
    }
}
static void b2_final(b2s *s, uint8_t out[32]) {
    s->t += s->buflen;
    memset(s->buf + s->buflen, 0, 64 - s->buflen);
    uint32_t m[16];
    memcpy(m, s->buf, 64);
    b2_compress(s->h, m, s->t, 1);
    memcpy(out, s->h, 32);
}
void ora_blake2s(const uint8_t *data, size_t len, uint8_t out[32]) {
    b2s s;
    b2_init(&s);
    b2_update(&s, data, len);
    b2_final(&s, out);
}

/* ======================================================================
 * Lifted Blake2s Merkle (src/prover/vcs_lifted/{prover,columns,leaves,layers}.zig,
 * src/core/vcs_lifted/blake2_merkle.zig:9-262)
 * leaf[p] = H("leaf"+0^60 || LE32(col[lift(p)]) for cols sorted by (log asc, index))
 * node    = H("node"+0^60 || left || right)
 * ==================================================================== */
static uint8_t LEAF_PREFIX[64] = {'l', 'e', 'a', 'f'};
static uint8_t NODE_PREFIX[64] = {'n', 'o', 'd', 'e'};

struct ora_tree {
    uint32_t max_log;
    uint8_t **layers; /* layers[k]: 2^k hashes, root first (prover.zig:23) */
};

typedef struct { const uint32_t *values; uint32_t log_size; uint32_t index; } colref;
static int colref_cmp(const void *a, const void *b) {
    const colref *x = a, *y = b;
    if (x->log_size != y->log_size) return x->log_size < y->log_size ? -1 : 1;
    return x->index < y->index ? -1 : (x->index > y->index);
}
/* Lifting index (leaves.zig:105-111; quotient_column_geometry.zig:38-42). */
static inline uint64_t lift_index(uint64_t pos, uint32_t max_log, uint32_t log_size) {
    uint32_t sh = max_log - log_size + 1;
    return ((pos >> sh) << 1) + (pos & 1);
}

ora_tree *ora_merkle_commit(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols) {
    colref *sorted = malloc(sizeof(colref) * (n_cols ? n_cols : 1));
    for (uint32_t i = 0; i < n_cols; i++) sorted[i] = (colref){cols[i], log_sizes[i], i};
    qsort(sorted, n_cols, sizeof(colref), colref_cmp); /* stable: ties broken by index */
    uint32_t max_log = n_cols ? sorted[n_cols - 1].log_size : 0;
    ora_tree *t = malloc(sizeof(ora_tree));
    t->max_log = max_log;
    t->layers = calloc(max_log + 1, sizeof(uint8_t *));
    uint64_t n_leaves = 1ull << max_log;
    uint8_t *leaves = malloc(n_leaves * 32);
    b2s seed;
    b2_init(&seed);
    b2_update(&seed, LEAF_PREFIX, 64);
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < (int64_t)n_leaves; p++) {
        b2s s = seed;
        for (uint32_t c = 0; c < n_cols; c++) {
            uint32_t v = sorted[c].values[lift_index((uint64_t)p, max_log, sorted[c].log_size)];
            b2_update(&s, (const uint8_t *)&v, 4);
        }
        b2_final(&s, leaves + 32 * p);
    }
    t->layers[max_log] = leaves;
    b2s nseed;
    b2_init(&nseed);
    b2_update(&nseed, NODE_PREFIX, 64);
    for (int32_t k = (int32_t)max_log - 1; k >= 0; k--) {
        uint64_t n = 1ull << k;
        uint8_t *layer = malloc(n * 32);
        const uint8_t *child = t->layers[k + 1];
#pragma omp parallel for schedule(static) if (n > 4096)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            b2s s = nseed;
            b2_update(&s, child + 64 * i, 64);
            b2_final(&s, layer + 32 * i);
        }
        t->layers[k] = layer;
    }
    free(sorted);
    return t;
}
void ora_tree_root(const ora_tree *t, uint8_t out[32]) { memcpy(out, t->layers[0], 32); }
uint32_t ora_tree_max_log(const ora_tree *t) { return t->max_log; }
void ora_tree_layer(const ora_tree *t, uint32_t layer_log, uint8_t *out) {
    memcpy(out, t->layers[layer_log], 32ull << layer_log);
}
void ora_tree_free(ora_tree *t) {
    if (!t) return;
    for (uint32_t k = 0; k <= t->max_log; k++) free(t->layers[k]);
    free(t->layers);
    free(t);
}

/* ======================================================================
 * Growable byte/word buffers for proof assembly.
 * ==================================================================== */
typedef struct { uint8_t *d; size_t n, cap; } bytes;
static void by_push(bytes *b, const void *src, size_t n) {
    if (b->n + n > b->cap) {
        b->cap = (b->n + n) * 2 + 64;
        b->d = realloc(b->d, b->cap);
    }
    memcpy(b->d + b->n, src, n);
    b->n += n;
}
static void by_str(bytes *b, const char *s) { by_push(b, s, strlen(s)); }
static void by_u64(bytes *b, uint64_t v) {
    char tmp[32];
    snprintf(tmp, sizeof tmp, "%llu", (unsigned long long)v);
    by_str(b, tmp);
}

/* Decommitment (src/prover/vcs_lifted/decommit.zig:56-237): hash witness for
 * sorted unique positions, queried values per column at lifted positions. */
typedef struct { bytes witness; /* 32-byte hashes */ } decommitment;
static void merkle_decommit_witness(const ora_tree *t, const uint64_t *pos, size_t npos, bytes *witness) {
    uint64_t *prev = malloc(sizeof(uint64_t) * (npos + 1)), *cur = malloc(sizeof(uint64_t) * (npos + 1));
    size_t nprev = 0;
    for (size_t i = 0; i < npos; i++)
        if (i == 0 || pos[i - 1] != pos[i]) prev[nprev++] = pos[i];
    for (int32_t layer = (int32_t)t->max_log - 1; layer >= 0; layer--) {
        const uint8_t *child = t->layers[layer + 1];
        size_t ncur = 0, q = 0;
        while (q < nprev) {
            uint64_t first = prev[q];
            int has_sib = q + 1 < nprev && (first ^ 1) == prev[q + 1];
            if (!has_sib) by_push(witness, child + 32 * (first ^ 1), 32);
            cur[ncur++] = first >> 1;
            q += has_sib ? 2 : 1;
        }
        uint64_t *tmp = prev; prev = cur; cur = tmp;
        nprev = ncur;
    }
    free(prev);
    free(cur);
}

/* ======================================================================
 * Blake2s channel (src/core/channel/blake2s.zig:32-287; mixRoot blake2_merkle.zig:237-249)
 * ==================================================================== */
typedef struct { uint8_t digest[32]; uint32_t n_draws; } channel;
static void ch_update(channel *ch, const uint8_t d[32]) {
    memcpy(ch->digest, d, 32);
    ch->n_draws = 0;
}
static void ch_mix_bytes(channel *ch, const void *data, size_t n) {
    b2s s;
    b2_init(&s);
    b2_update(&s, ch->digest, 32);
    b2_update(&s, data, n);
    uint8_t out[32];
    b2_final(&s, out);
    ch_update(ch, out);
}
static void ch_mix_felts(channel *ch, const qm31 *f, size_t n) {
    uint32_t *w = malloc(16 * (n ? n : 1));
    for (size_t i = 0; i < n; i++) q_store(f[i], w + 4 * i);
    ch_mix_bytes(ch, w, 16 * n);
    free(w);
}
static void ch_mix_u32s(channel *ch, const uint32_t *d, size_t n) { ch_mix_bytes(ch, d, 4 * n); }
static void ch_mix_u64(channel *ch, uint64_t v) {
    uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
    ch_mix_u32s(ch, w, 2);
}
static void ch_mix_root(channel *ch, const uint8_t root[32]) { ch_mix_bytes(ch, root, 32); }
static void ch_draw_u32s(channel *ch, uint32_t out[8]) {
    uint8_t in[37];
    memcpy(in, ch->digest, 32);
    memcpy(in + 32, &ch->n_draws, 4);
    in[36] = 0;
    ch->n_draws++;
    uint8_t h[32];
    ora_blake2s(in, 37, h);
    memcpy(out, h, 32);
}
static void ch_draw_base_felts(channel *ch, uint32_t out[8]) {
    for (;;) {
        uint32_t w[8];
        ch_draw_u32s(ch, w);
        int ok = 1;
        for (int i = 0; i < 8; i++)
            if (w[i] >= 2 * P) ok = 0;
        if (!ok) continue;
        for (int i = 0; i < 8; i++) out[i] = m_from_u64(w[i]);
        return;
    }
}
static qm31 ch_draw_secure_felt(channel *ch) {
    uint32_t f[8];
    ch_draw_base_felts(ch, f);
    return q_load(f);
}
/* PoW (blake2s.zig:135-166,289-300). */
static int pow_check(const uint8_t prefix[32], uint64_t nonce, uint32_t bits) {
    uint8_t in[40], h[32];
    memcpy(in, prefix, 32);
    memcpy(in + 32, &nonce, 8);
    ora_blake2s(in, 40, h);
    uint32_t count = 0;
    for (int i = 0; i < 16; i++) {
        if (h[i] == 0) { count += 8; continue; }
        count += __builtin_ctz(h[i]);
        break;
    }
    return count >= bits;
}
uint64_t ora_pow_grind(const uint8_t digest[32], uint32_t pow_bits) {
    if (pow_bits == 0) return 0;
    uint8_t in[52], prefix[32];
    uint32_t pp = 0x12345678u;
    memcpy(in, &pp, 4);
    memset(in + 4, 0, 12);
    memcpy(in + 16, digest, 32);
    memcpy(in + 48, &pow_bits, 4);
    ora_blake2s(in, 52, prefix);
    /* Lowest valid nonce: search ascending chunks in parallel (result is
     * scheduling independent, as in grindWorker's fetchMin). */
    const uint64_t chunk = 1u << 16;
    for (uint64_t base = 0;; base += chunk) {
        uint64_t best = UINT64_MAX;
#pragma omp parallel for schedule(static) reduction(min : best)
        for (int64_t i = 0; i < (int64_t)chunk; i++)
            if (pow_check(prefix, base + (uint64_t)i, pow_bits) && base + (uint64_t)i < best) best = base + (uint64_t)i;
        if (best != UINT64_MAX) return best;
    }
}

/* ======================================================================
 * FRI folds (src/core/fri/folding.zig:316-380,558-602)
 * ==================================================================== */
void ora_fold_line(const uint32_t *src, uint32_t log_src, uint32_t coset_initial_index, const uint32_t alpha[4],
                   uint32_t *dst) {
    coset c = coset_new(coset_initial_index, log_src);
    qm31 a = q_load(alpha);
    uint64_t n = 1ull << (log_src - 1);
    uint32_t lg = log_src - 1;
#pragma omp parallel for schedule(static) if (n > 4096)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t x = coset_at(c, bitrev((uint64_t)i, lg)).x;
        qm31 f0 = q_load(src + 8 * i), f1 = q_load(src + 8 * i + 4);
        qm31 s = q_add(f0, f1), d = q_mul_m(q_sub(f0, f1), m_inv(x));
        q_store(q_add(s, q_mul(a, d)), dst + 4 * i);
    }
}
void ora_fold_circle_into_line(const uint32_t *const src[4], uint32_t log_src, const uint32_t alpha[4],
                               uint32_t *dst) {
    coset hc = coset_half_odds(log_src - 1);
    qm31 a = q_load(alpha), a2 = q_square(q_load(alpha));
    uint64_t n = 1ull << (log_src - 1);
    uint32_t lg = log_src - 1;
#pragma omp parallel for schedule(static) if (n > 4096)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t y = coset_at(hc, bitrev((uint64_t)i, lg)).y;
        qm31 f0 = {{src[0][2 * i], src[1][2 * i]}, {src[2][2 * i], src[3][2 * i]}};
        qm31 f1 = {{src[0][2 * i + 1], src[1][2 * i + 1]}, {src[2][2 * i + 1], src[3][2 * i + 1]}};
        qm31 s = q_add(f0, f1), d = q_mul_m(q_sub(f0, f1), m_inv(y));
        qm31 fp = q_add(s, q_mul(a, d));
        qm31 cur = q_load(dst + 4 * i);
        q_store(q_add(q_mul(cur, a2), fp), dst + 4 * i);
    }
}

/* ======================================================================
 * Wide Fibonacci trace (src/examples/wide_fibonacci/trace.zig:40-162)
 * ==================================================================== */
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
/* Seeded uniform M31 word k of a trace (mode 1): splitmix64(seed ^ k*golden) mod p. */
static inline uint32_t seeded_m31(uint64_t seed, uint64_t k) {
    return (uint32_t)(splitmix64(seed ^ (k * 0xD1B54A32D192ED03ull)) % P);
}
void ora_widefib_trace(uint32_t log_n, uint32_t seq_len, uint32_t mode, uint64_t seed, uint32_t *out) {
    uint64_t n = 1ull << log_n;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)n; s++) {
        uint32_t a, b;
        if (mode == 0) {
            /* logicalRow = circleDomainIndexToCosetIndex(bitReverse(s)) (trace.zig:159-162; utils.zig) */
            uint64_t ci = bitrev((uint64_t)s, log_n);
            uint64_t row = ci < n / 2 ? ci * 2 : (n - 1 - ci) * 2 + 1;
            a = 1;
            b = (uint32_t)row;
        } else {
            a = seeded_m31(seed, 2 * (uint64_t)s);
            b = seeded_m31(seed, 2 * (uint64_t)s + 1);
        }
        out[s] = a;
        out[n + s] = b;
        for (uint32_t c = 2; c < seq_len; c++) {
            uint32_t nx = m_add(m_mul(a, a), m_mul(b, b));
            out[(uint64_t)c * n + s] = nx;
            a = b;
            b = nx;
        }
    }
}

/* ======================================================================
 * Poseidon AIR trace (src/examples/poseidon/input.zig:9-211): 8 permutation instances per
 * row, each of 16 state words through 4 full rounds, 14 partial rounds and 4 full rounds;
 * the columns of one instance are the input state, the state after every full round and
 * state[0] after every partial round (158 columns per instance, 1264 per row).
 * ==================================================================== */
#define POS_STATE 16
#define POS_PARTIAL 14
#define POS_HALF_FULL 4
#define POS_COLS_PER_REP (POS_STATE * (1 + 2 * POS_HALF_FULL) + POS_PARTIAL)
#define POS_COLS (POS_COLS_PER_REP * 8)
static inline uint32_t pos_pow5(uint32_t x) {
    uint32_t x2 = m_mul(x, x);
    return m_mul(m_mul(x2, x2), x);
}
/* M4 of the external matrix (input.zig:170-183), as a linear map on four words. */
static void pos_m4(uint32_t *x) {
    uint32_t t0 = m_add(x[0], x[1]), t02 = m_add(t0, t0);
    uint32_t t1 = m_add(x[2], x[3]), t12 = m_add(t1, t1);
    uint32_t t2 = m_add(m_add(x[1], x[1]), t1);
    uint32_t t3 = m_add(m_add(x[3], x[3]), t0);
    uint32_t t4 = m_add(m_add(t12, t12), t3);
    uint32_t t5 = m_add(m_add(t02, t02), t2);
    x[0] = m_add(t3, t5);
    x[1] = t5;
    x[2] = m_add(t2, t4);
    x[3] = t4;
}
static void pos_external_round(uint32_t *s, uint32_t round) {
    for (int i = 0; i < POS_STATE; i++) s[i] = m_add(s[i], (uint32_t)((1234u + round * 37u + (uint32_t)i) % P));
    for (int i = 0; i < 4; i++) pos_m4(s + 4 * i);
    for (int j = 0; j < 4; j++) {
        uint32_t sum = m_add(m_add(s[j], s[j + 4]), m_add(s[j + 8], s[j + 12]));
        for (int i = 0; i < 4; i++) s[4 * i + j] = m_add(s[4 * i + j], sum);
    }
    for (int i = 0; i < POS_STATE; i++) s[i] = pos_pow5(s[i]);
}
static void pos_internal_round(uint32_t *s, uint32_t round) {
    s[0] = m_add(s[0], (uint32_t)((9876u + round * 17u) % P));
    uint32_t sum = 0;
    for (int i = 0; i < POS_STATE; i++) sum = m_add(sum, s[i]);
    for (int i = 0; i < POS_STATE; i++) s[i] = m_add(m_mul(s[i], (uint32_t)((1ull << (i + 1)) % P)), sum);
    s[0] = pos_pow5(s[0]);
}
void ora_poseidon_trace(uint32_t log_n_rows, uint32_t *out) {
    uint64_t n = 1ull << log_n_rows;
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < (int64_t)n; row++) {
        uint64_t col = 0;
        for (uint32_t rep = 0; rep < 8; rep++) {
            uint32_t s[POS_STATE];
            for (int i = 0; i < POS_STATE; i++) {
                s[i] = (uint32_t)(((uint64_t)row * POS_STATE + (uint64_t)i + rep) % P);
                out[col++ * n + row] = s[i];
            }
            for (uint32_t r = 0; r < POS_HALF_FULL; r++) {
                pos_external_round(s, r);
                for (int i = 0; i < POS_STATE; i++) out[col++ * n + row] = s[i];
            }
            for (uint32_t r = 0; r < POS_PARTIAL; r++) {
                pos_internal_round(s, r);
                out[col++ * n + row] = s[0];
            }
            for (uint32_t r = 0; r < POS_HALF_FULL; r++) {
                pos_external_round(s, r + POS_HALF_FULL);
                for (int i = 0; i < POS_STATE; i++) out[col++ * n + row] = s[i];
            }
        }
    }
}

/* ======================================================================
 * Periodic XOR example AIR (src/examples/xor/input.zig:31-107; core/air/utils.zig:23-65):
 * preprocessed IsFirst and IsStepWithOffset indicators and one main column, each in
 * bit-reversed circle-domain order.
 * ==================================================================== */
static uint64_t circle_bitrev_index(uint32_t log_size, uint64_t coset_index) {
    uint64_t n = 1ull << log_size;
    uint64_t ci = (coset_index & 1) == 0 ? coset_index / 2 : (2 * n - coset_index) / 2;
    return bitrev(ci, log_size);
}
void ora_xor_example_columns(uint32_t log_size, uint32_t log_step, uint64_t offset, uint32_t *is_first,
                             uint32_t *is_step, uint32_t *main_col) {
    uint64_t n = 1ull << log_size, step = 1ull << log_step;
    memset(is_first, 0, 4 * n);
    memset(is_step, 0, 4 * n);
    is_first[0] = 1;
    for (uint64_t i = offset % step; i < n; i += step) is_step[circle_bitrev_index(log_size, i)] = 1;
    for (uint64_t i = 0; i < n; i++) main_col[circle_bitrev_index(log_size, i)] = (i & 1) == 0 ? 1 : 0;
}

/* ======================================================================
 * Proof assembly
 * ==================================================================== */
static double g_timings[8];
void ora_last_timings(double out[8]) { memcpy(out, g_timings, sizeof g_timings); }
/* Thread count of the OpenMP regions (the CPU baseline is timed at all host threads and at 1). */
void ora_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int ora_get_threads(void) { return omp_get_max_threads(); }
static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

typedef struct {
    uint32_t n_cols;
    uint32_t *log_sizes;  /* LDE log sizes */
    uint32_t **lde;       /* evaluations on CanonicCoset(log).circleDomain(), bitrev */
    uint32_t **coeffs;    /* retained coefficients (log - blowup) */
    uint32_t *coeff_logs;
    ora_tree *tree;
} ctree;

static void ctree_free(ctree *t) {
    for (uint32_t i = 0; i < t->n_cols; i++) {
        free(t->lde[i]);
        free(t->coeffs[i]);
    }
    free(t->lde);
    free(t->coeffs);
    free(t->log_sizes);
    free(t->coeff_logs);
    ora_tree_free(t->tree);
}

/* commitPolys / commit (pcs/scheme.zig:115-140; commit_polys.zig:10-67): coefficients ->
 * LDE (blowup) -> Merkle -> mixRoot.  Takes ownership of coeffs. */
static void commit_coeffs(ctree *t, uint32_t n_cols, uint32_t **coeffs, const uint32_t *coeff_logs, uint32_t blowup,
                          channel *ch) {
    t->n_cols = n_cols;
    t->coeffs = coeffs;
    t->coeff_logs = malloc(4 * (n_cols ? n_cols : 1));
    t->log_sizes = malloc(4 * (n_cols ? n_cols : 1));
    t->lde = malloc(sizeof(uint32_t *) * (n_cols ? n_cols : 1));
    for (uint32_t i = 0; i < n_cols; i++) {
        t->coeff_logs[i] = coeff_logs[i];
        t->log_sizes[i] = coeff_logs[i] + blowup;
        t->lde[i] = malloc(4ull << t->log_sizes[i]);
    }
    for (uint32_t i = 0; i < n_cols; i++) get_twiddles(t->log_sizes[i] > 2 ? t->log_sizes[i] : 3);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)n_cols; i++) ora_circle_lde(coeffs[i], coeff_logs[i], blowup, t->lde[i]);
    t->tree = ora_merkle_commit((const uint32_t *const *)t->lde, t->log_sizes, n_cols);
    uint8_t root[32];
    ora_tree_root(t->tree, root);
    ch_mix_root(ch, root);
}

static void json_hash(bytes *b, const uint8_t *h) {
    by_str(b, "[");
    for (int i = 0; i < 32; i++) {
        if (i) by_str(b, ",");
        by_u64(b, h[i]);
    }
    by_str(b, "]");
}
static void json_qm31(bytes *b, qm31 v) {
    uint32_t w[4];
    q_store(v, w);
    by_str(b, "[");
    for (int i = 0; i < 4; i++) {
        if (i) by_str(b, ",");
        by_u64(b, w[i]);
    }
    by_str(b, "]");
}
static void json_hash_list(bytes *b, const bytes *w) {
    by_str(b, "[");
    for (size_t i = 0; i < w->n / 32; i++) {
        if (i) by_str(b, ",");
        json_hash(b, w->d + 32 * i);
    }
    by_str(b, "]");
}

typedef struct {
    bytes json_layer; /* encoded FriLayerWire */
} fri_layer_out;

/* decommitLayerExtended for fold_step 1 (prover/fri.zig:626-740): pairs of
 * positions, witness = non-queried partner, Merkle decommit at the pairs. */
static void fri_layer_json(bytes *out, const ora_tree *tree, const qm31 *col, const uint64_t *q, size_t nq) {
    uint64_t *dpos = malloc(sizeof(uint64_t) * (2 * nq + 1));
    size_t nd = 0;
    qm31 *wit = malloc(sizeof(qm31) * (2 * nq + 1));
    size_t nw = 0;
    size_t i = 0;
    while (i < nq) {
        uint64_t key = q[i] >> 1;
        size_t j = i + 1;
        while (j < nq && (q[j] >> 1) == key) j++;
        size_t at = i;
        for (uint64_t p = key << 1; p < (key << 1) + 2; p++) {
            dpos[nd++] = p;
            if (at < j && q[at] == p) at++;
            else wit[nw++] = col[p];
        }
        i = j;
    }
    bytes hw = {0};
    merkle_decommit_witness(tree, dpos, nd, &hw);
    uint8_t root[32];
    ora_tree_root(tree, root);
    by_str(out, "{\"fri_witness\":[");
    for (size_t k = 0; k < nw; k++) {
        if (k) by_str(out, ",");
        json_qm31(out, wit[k]);
    }
    by_str(out, "],\"decommitment\":{\"hash_witness\":");
    json_hash_list(out, &hw);
    by_str(out, "},\"commitment\":");
    json_hash(out, root);
    by_str(out, "}");
    free(hw.d);
    free(dpos);
    free(wit);
}

static ora_tree *commit_qm31_column(const qm31 *v, uint32_t log_size, uint32_t **coords_out) {
    uint64_t n = 1ull << log_size;
    uint32_t *c[4];
    for (int k = 0; k < 4; k++) c[k] = malloc(4 * n);
    for (uint64_t i = 0; i < n; i++) {
        c[0][i] = v[i].c0.a; c[1][i] = v[i].c0.b; c[2][i] = v[i].c1.a; c[3][i] = v[i].c1.b;
    }
    uint32_t logs[4] = {log_size, log_size, log_size, log_size};
    ora_tree *t = ora_merkle_commit((const uint32_t *const *)c, logs, 4);
    if (coords_out)
        for (int k = 0; k < 4; k++) coords_out[k] = c[k];
    else
        for (int k = 0; k < 4; k++) free(c[k]);
    return t;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

/* The AIRs this proof core serves: the wide-Fibonacci component and the Poseidon
 * component (src/examples/poseidon.zig:397-504), whose composition is the constant
 * (log_n_instances, 158, 1264, 1) (poseidon.zig:562-569). */
enum { AIR_WIDEFIB = 0, AIR_POSEIDON = 1, AIR_XOR = 2 };

/* air AIR_XOR: the periodic XOR example (src/examples/xor.zig:419-605): two preprocessed
 * columns and one main column, statement (log_size, log_step, offset), constant
 * composition (log_size, log_step, offset, 1); xs holds log_step and offset. */
/* randomSecureFieldPoint (circle.zig:159-167). */
static qpt draw_oods_point(channel *ch) {
    qm31 tt = ch_draw_secure_felt(ch);
    qm31 t2_ = q_square(tt);
    qm31 inv1 = q_inv(q_add(t2_, Q_ONE));
    return (qpt){q_mul(q_sub(Q_ONE, t2_), inv1), q_mul(q_add(tt, tt), inv1)};
}
static inline int qp_eq(qpt a, qpt b) { return q_eq(a.x, b.x) && q_eq(a.y, b.y); }
/* QM31 circle point plus an M31 circle point. */
static inline qpt qp_add_m(qpt a, pt b) {
    qm31 bx = q_from_m(b.x), by = q_from_m(b.y);
    return (qpt){q_sub(q_mul(a.x, bx), q_mul(a.y, by)), q_add(q_mul(a.x, by), q_mul(a.y, bx))};
}

/* The rest of the proof from the committed trees (the composition tree last): sampled values
 * at the mask points (sampled_values.zig), their transcript, the quotient column with one batch
 * per distinct point (core/pcs/quotients/samples.zig:245-312 with the periodicity sample of
 * two-point columns, row_evaluation.zig:402-521), FRI (prover/fri.zig:163-621), PoW
 * (scheme.zig:733-735), queries (core/queries.zig:8-60), decommitments and the proof wire
 * (interop/proof_wire.zig:19-58).  ns[col] (1 or 2) and pts[2*col + s]: the mask points of every
 * column in tree/column order. */
static int prove_tail(channel *chp, ctree *trees, int ntr, const uint32_t *ns, const qpt *pts,
                      const ora_widefib_params *prm, const double tm[3], double t3, uint8_t *out, size_t cap,
                      size_t *out_len) {
    channel ch = *chp;
    const uint32_t B = prm->log_blowup_factor;
    uint32_t L = 0, total_cols = 0;
    for (int tr = 0; tr < ntr; tr++) {
        total_cols += trees[tr].n_cols;
        for (uint32_t c = 0; c < trees[tr].n_cols; c++)
            if (trees[tr].log_sizes[c] > L) L = trees[tr].log_sizes[c];
    }
    /* flat column tables */
    const uint32_t **fl_lde = malloc(sizeof(uint32_t *) * total_cols);
    const uint32_t **fl_coef = malloc(sizeof(uint32_t *) * total_cols);
    uint32_t *fl_log = malloc(4 * total_cols), *fl_clog = malloc(4 * total_cols), *first_s = malloc(4 * (total_cols + 1));
    uint32_t n_samples = 0;
    {
        uint32_t at = 0;
        for (int tr = 0; tr < ntr; tr++)
            for (uint32_t c = 0; c < trees[tr].n_cols; c++, at++) {
                fl_lde[at] = trees[tr].lde[c];
                fl_coef[at] = trees[tr].coeffs[c];
                fl_log[at] = trees[tr].log_sizes[c];
                fl_clog[at] = trees[tr].coeff_logs[c];
                first_s[at] = n_samples;
                n_samples += ns[at];
            }
        first_s[total_cols] = n_samples;
    }
    /* Sampled values: each column at its mask points folded (L - log) times (the lifted PCS). */
    qm31 *sv = malloc(sizeof(qm31) * (n_samples ? n_samples : 1));
    qpt *spt = malloc(sizeof(qpt) * (n_samples ? n_samples : 1));
    for (uint32_t i = 0; i < total_cols; i++)
        for (uint32_t k = 0; k < ns[i]; k++) {
            qpt p = pts[2 * i + k];
            for (uint32_t f = fl_log[i]; f < L; f++) { /* repeatedDouble */
                qm31 xx = q_square(p.x), xy = q_mul(p.x, p.y);
                p.x = q_sub(q_add(xx, xx), Q_ONE);
                p.y = q_add(xy, xy);
            }
            spt[first_s[i] + k] = p;
        }
    {
        /* one weight vector per distinct (point, coefficient log) */
        enum { MAXW = 16 };
        qpt wpt[MAXW];
        uint32_t wlog[MAXW];
        qm31 *wv[MAXW];
        int nw = 0;
        for (uint32_t i = 0; i < total_cols; i++)
            for (uint32_t k = 0; k < ns[i]; k++) {
                qpt p = spt[first_s[i] + k];
                int w = 0;
                for (; w < nw; w++)
                    if (wlog[w] == fl_clog[i] && qp_eq(wpt[w], p)) break;
                if (w == nw) {
                    if (nw == MAXW) return 4;
                    wpt[nw] = p;
                    wlog[nw] = fl_clog[i];
                    wv[nw] = eval_weights(p, fl_clog[i]);
                    nw++;
                }
            }
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t i = 0; i < (int64_t)total_cols; i++)
            for (uint32_t k = 0; k < ns[i]; k++) {
                qpt p = spt[first_s[i] + k];
                int w = 0;
                while (!(wlog[w] == fl_clog[i] && qp_eq(wpt[w], p))) w++;
                sv[first_s[i] + k] = fl_clog[i] == 0 ? q_from_m(fl_coef[i][0])
                                                      : dot_weights(fl_coef[i], wv[w], 1ull << fl_clog[i]);
            }
        for (int w = 0; w < nw; w++) free(wv[w]);
    }
    /* sampled_value_transcript.mixIntoChannel: tree, column, sample order. */
    ch_mix_felts(&ch, sv, n_samples);
    double t4 = now_ms();

    /* Quotients (core/pcs/quotients/{samples,row_evaluation}.zig; constraints.zig:87-97): one batch
     * per distinct point in first-seen order; a two-point column first adds its second sample at
     * that point plus the lifting domain's step doubled log times (samples.zig:270-285); random
     * coefficients 1, q, q^2, ... in that order. */
    qm31 qalpha = ch_draw_secure_felt(&ch);
    uint64_t LN = 1ull << L;
    uint32_t max_e = n_samples + total_cols;
    qpt *bpt = malloc(sizeof(qpt) * max_e);
    qm31 *bsa = malloc(sizeof(qm31) * max_e), *bsb = malloc(sizeof(qm31) * max_e);
    uint32_t *ecol = malloc(4 * max_e), *ebat = malloc(4 * max_e);
    qm31 *ecoef = malloc(sizeof(qm31) * max_e);
    uint32_t nb = 0, ne = 0;
    {
        qm31 rp = Q_ONE;
        const uint32_t lift_gen = subgroup_gen(L);
        for (uint32_t i = 0; i < total_cols; i++) {
            for (int32_t k = ns[i] == 2 ? -1 : 0; k < (int32_t)ns[i]; k++) {
                qpt p;
                qm31 v;
                if (k < 0) {
                    p = qp_add_m(spt[first_s[i] + 1],
                                 idx_to_point((uint32_t)(((uint64_t)lift_gen << fl_log[i]) & ORDER_MASK)));
                    v = sv[first_s[i] + 1];
                } else {
                    p = spt[first_s[i] + k];
                    v = sv[first_s[i] + k];
                }
                uint32_t b = 0;
                for (; b < nb; b++)
                    if (qp_eq(bpt[b], p)) break;
                if (b == nb) {
                    bpt[nb] = p;
                    bsa[nb] = Q_ZERO;
                    bsb[nb] = Q_ZERO;
                    nb++;
                }
                qm31 a = q_sub(q_conj(v), v);
                qm31 c = q_sub(q_conj(p.y), p.y);
                qm31 bb = q_sub(q_mul(v, c), q_mul(a, p.y));
                ecol[ne] = i;
                ebat[ne] = b;
                ecoef[ne] = q_mul(rp, c);
                ne++;
                bsa[b] = q_add(bsa[b], q_mul(rp, a));
                bsb[b] = q_add(bsb[b], q_mul(rp, bb));
                rp = q_mul(rp, qalpha);
            }
        }
    }
    qm31 *quot = malloc(sizeof(qm31) * LN);
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < (int64_t)LN; row++) {
        pt dp = domain_at(L, bitrev((uint64_t)row, L));
        qm31 acc = Q_ZERO;
        for (uint32_t b = 0; b < nb; b++) {
            qm31 num = Q_ZERO;
            for (uint32_t e = 0; e < ne; e++) {
                if (ebat[e] != b) continue;
                uint32_t i = ecol[e];
                uint32_t v = fl_lde[i][lift_index((uint64_t)row, L, fl_log[i])];
                num = q_add(num, q_mul_m(ecoef[e], v));
            }
            num = q_sub(num, q_add(q_mul_m(bsa[b], dp.y), bsb[b]));
            cm31 prx = bpt[b].x.c0, pry = bpt[b].y.c0, pix = bpt[b].x.c1, piy = bpt[b].y.c1;
            cm31 det = c_sub(c_mul(prx, piy), c_mul(pry, pix));
            cm31 den = c_add(c_sub(det, c_mul_m(piy, dp.x)), c_mul_m(pix, dp.y));
            acc = q_add(acc, q_mul_c(num, c_inv(den)));
        }
        quot[row] = acc;
    }
    free(bpt);
    free(bsa);
    free(bsb);
    free(ecol);
    free(ebat);
    free(ecoef);
    double t5 = now_ms();

    /* FRI commit (prover/fri.zig:163-621). */
    uint32_t nlayers_max = L + 1;
    ora_tree **ftrees = calloc(nlayers_max, sizeof(ora_tree *));
    qm31 **fcols = calloc(nlayers_max, sizeof(qm31 *));
    uint32_t *flogs = calloc(nlayers_max, 4);
    uint32_t *first_coords[4];
    ftrees[0] = commit_qm31_column(quot, L, first_coords);
    fcols[0] = quot;
    flogs[0] = L;
    {
        uint8_t r[32];
        ora_tree_root(ftrees[0], r);
        ch_mix_root(&ch, r);
    }
    qm31 fa = ch_draw_secure_felt(&ch);
    uint64_t ln = LN / 2;
    qm31 *line = calloc(ln, sizeof(qm31));
    ora_fold_circle_into_line((const uint32_t *const *)first_coords, L, (const uint32_t *)&fa, (uint32_t *)line);
    for (int k = 0; k < 4; k++) free(first_coords[k]);
    coset ldom = coset_half_odds(L - 1);
    uint32_t llog = L - 1;
    const uint32_t last_log = prm->log_last_layer_degree_bound + B;
    uint32_t ninner = 0;
    while (llog > last_log) {
        ftrees[1 + ninner] = commit_qm31_column(line, llog, NULL);
        fcols[1 + ninner] = line;
        flogs[1 + ninner] = llog;
        uint8_t r[32];
        ora_tree_root(ftrees[1 + ninner], r);
        ch_mix_root(&ch, r);
        qm31 la = ch_draw_secure_felt(&ch);
        qm31 *next = malloc(sizeof(qm31) * (1ull << (llog - 1)));
        ora_fold_line((const uint32_t *)line, llog, ldom.initial_index, (const uint32_t *)&la, (uint32_t *)next);
        ldom = coset_double(ldom);
        llog--;
        line = next;
        ninner++;
    }
    /* Last layer (fri.zig:594-621; line_evaluation.zig:97-151; line.zig:134-142). */
    uint64_t lsz = 1ull << llog;
    bitrev_q(line, lsz);
    {
        coset cd = ldom;
        while ((1ull << cd.log_size) > 1) {
            uint64_t csz = 1ull << cd.log_size, half = csz / 2;
            for (uint64_t st0 = 0; st0 < lsz; st0 += csz) {
                pt cur = idx_to_point(cd.initial_index), step = idx_to_point(cd.step_size);
                for (uint64_t i = 0; i < half; i++) {
                    uint32_t ix = m_inv(cur.x);
                    qm31 v0 = line[st0 + i], v1 = line[st0 + half + i];
                    line[st0 + i] = q_add(v0, v1);
                    line[st0 + half + i] = q_mul_m(q_sub(v0, v1), ix);
                    cur = p_add(cur, step);
                }
            }
            cd = coset_double(cd);
        }
        uint32_t li = m_inv((uint32_t)(lsz % P));
        for (uint64_t i = 0; i < lsz; i++) line[i] = q_mul_m(line[i], li);
    }
    bitrev_q(line, lsz); /* ordered coefficients */
    uint64_t bound = 1ull << prm->log_last_layer_degree_bound;
    for (uint64_t i = bound; i < lsz; i++)
        if (!q_is_zero(line[i])) return 3; /* InvalidLastLayerDegree */
    qm31 *lastp = malloc(sizeof(qm31) * bound);
    memcpy(lastp, line, sizeof(qm31) * bound);
    bitrev_q(lastp, bound); /* fromOrderedCoefficients */
    free(line);
    ch_mix_felts(&ch, lastp, bound);
    double t6 = now_ms();

    /* PoW (scheme.zig:733-735). */
    uint64_t nonce = ora_pow_grind(ch.digest, prm->pow_bits);
    ch_mix_u64(&ch, nonce);
    double t7 = now_ms();

    /* Queries (core/queries.zig:8-60). */
    uint32_t nq = prm->n_queries;
    uint64_t *qs = malloc(sizeof(uint64_t) * (nq + 8));
    {
        uint32_t got = 0;
        while (got < nq) {
            uint32_t w[8];
            ch_draw_u32s(&ch, w);
            for (int i = 0; i < 8 && got < nq; i++) qs[got++] = w[i] & ((1u << L) - 1);
        }
    }
    qsort(qs, nq, sizeof(uint64_t), cmp_u64);
    size_t nuq = 0;
    for (uint32_t i = 0; i < nq; i++)
        if (nuq == 0 || qs[nuq - 1] != qs[i]) qs[nuq++] = qs[i];

    /* Assemble the proof wire (interop/proof_wire.zig:19-58). */
    bytes j = {0};
    by_str(&j, "{\"config\":{\"pow_bits\":");
    by_u64(&j, prm->pow_bits);
    by_str(&j, ",\"fri_config\":{\"log_blowup_factor\":");
    by_u64(&j, B);
    by_str(&j, ",\"log_last_layer_degree_bound\":");
    by_u64(&j, prm->log_last_layer_degree_bound);
    by_str(&j, ",\"n_queries\":");
    by_u64(&j, nq);
    by_str(&j, ",\"fold_step\":1},\"lifting_log_size\":null},\"commitments\":[");
    for (int tr = 0; tr < ntr; tr++) {
        uint8_t r[32];
        ora_tree_root(trees[tr].tree, r);
        if (tr) by_str(&j, ",");
        json_hash(&j, r);
    }
    by_str(&j, "],\"sampled_values\":[");
    {
        size_t at = 0;
        for (int tr = 0; tr < ntr; tr++) {
            by_str(&j, tr ? ",[" : "[");
            for (uint32_t c = 0; c < trees[tr].n_cols; c++, at++) {
                if (c) by_str(&j, ",");
                by_str(&j, "[");
                for (uint32_t k = 0; k < ns[at]; k++) {
                    if (k) by_str(&j, ",");
                    json_qm31(&j, sv[first_s[at] + k]);
                }
                by_str(&j, "]");
            }
            by_str(&j, "]");
        }
    }
    by_str(&j, "],\"decommitments\":[");
    for (int tr = 0; tr < ntr; tr++) {
        bytes hw = {0};
        if (trees[tr].n_cols) merkle_decommit_witness(trees[tr].tree, qs, nuq, &hw);
        if (tr) by_str(&j, ",");
        by_str(&j, "{\"hash_witness\":");
        json_hash_list(&j, &hw);
        by_str(&j, "}");
        free(hw.d);
    }
    by_str(&j, "],\"queried_values\":[");
    for (int tr = 0; tr < ntr; tr++) {
        if (tr) by_str(&j, ",");
        by_str(&j, "[");
        for (uint32_t c = 0; c < trees[tr].n_cols; c++) {
            if (c) by_str(&j, ",");
            by_str(&j, "[");
            for (size_t q = 0; q < nuq; q++) {
                if (q) by_str(&j, ",");
                by_u64(&j, trees[tr].lde[c][lift_index(qs[q], trees[tr].tree->max_log, trees[tr].log_sizes[c])]);
            }
            by_str(&j, "]");
        }
        by_str(&j, "]");
    }
    by_str(&j, "],\"proof_of_work\":");
    by_u64(&j, nonce);
    by_str(&j, ",\"fri_proof\":{\"first_layer\":");
    /* First layer column is the quotient (AoS view). */
    fri_layer_json(&j, ftrees[0], fcols[0], qs, nuq);
    by_str(&j, ",\"inner_layers\":[");
    uint64_t *lq = malloc(sizeof(uint64_t) * (nuq + 1));
    size_t nlq = 0;
    for (size_t i = 0; i < nuq; i++) {
        uint64_t v = qs[i] >> 1;
        if (nlq == 0 || lq[nlq - 1] != v) lq[nlq++] = v;
    }
    for (uint32_t k = 0; k < ninner; k++) {
        if (k) by_str(&j, ",");
        fri_layer_json(&j, ftrees[1 + k], fcols[1 + k], lq, nlq);
        size_t m2 = 0;
        for (size_t i = 0; i < nlq; i++) {
            uint64_t v = lq[i] >> 1;
            if (m2 == 0 || lq[m2 - 1] != v) lq[m2++] = v;
        }
        nlq = m2;
    }
    by_str(&j, "],\"last_layer_poly\":[");
    for (uint64_t i = 0; i < bound; i++) {
        if (i) by_str(&j, ",");
        json_qm31(&j, lastp[i]);
    }
    by_str(&j, "]}}");
    double t8 = now_ms();

    g_timings[0] = tm[1] - tm[0]; /* trace generation */
    g_timings[1] = tm[2] - tm[1]; /* main commit */
    g_timings[2] = t3 - tm[2];    /* composition + commit */
    g_timings[3] = t4 - t3;       /* sampled values */
    g_timings[4] = t5 - t4;       /* quotients */
    g_timings[5] = t6 - t5;       /* FRI commit */
    g_timings[6] = t7 - t6;       /* PoW */
    g_timings[7] = t8 - tm[1];    /* prove total (excludes trace generation) */

    int rc = 0;
    *out_len = j.n;
    if (j.n > cap) rc = 2;
    else memcpy(out, j.d, j.n);
    free(j.d);
    free(lq);
    free(qs);
    free(lastp);
    for (uint32_t k = 0; k <= ninner; k++) {
        ora_tree_free(ftrees[k]);
        free(fcols[k]);
    }
    free(ftrees);
    free(fcols);
    free(flogs);
    free(sv);
    free(spt);
    free(fl_lde);
    free(fl_coef);
    free(fl_log);
    free(fl_clog);
    free(first_s);
    *chp = ch;
    return rc;
}

static int prove_core(int air, const ora_widefib_params *prm, const uint64_t *xs, uint8_t *out, size_t cap,
                      size_t *out_len) {
    const uint32_t n = prm->log_n_rows, W = prm->sequence_len, B = prm->log_blowup_factor;
    if (n < 1 || W < 1 || B < 1 || (air == AIR_WIDEFIB && W < 3)) return 1;
    double t0 = now_ms();
    memset(g_timings, 0, sizeof g_timings);
    init_gen_doubles();
    channel ch;
    memset(&ch, 0, sizeof ch);
    /* PcsConfig.mixInto (core/pcs/mod.zig:32-52), fold_step 1, no lifting override. */
    qm31 cfg = {{prm->pow_bits, B}, {prm->n_queries, prm->log_last_layer_degree_bound}};
    ch_mix_felts(&ch, &cfg, 1);

    /* Tree 0: preprocessed; no columns for wide-Fib and Poseidon (examples/wide_fibonacci/
     * trace.zig:184-186), IsFirst and IsStepWithOffset for XOR (xor/input.zig:48-71). */
    ctree trees[3];
    memset(trees, 0, sizeof trees);
    uint64_t N = 1ull << n;
    uint32_t *trace = malloc(4ull * N * (W + 2));
    if (air == AIR_XOR) {
        ora_xor_example_columns(n, (uint32_t)xs[0], xs[1], trace + N * W, trace + N * (W + 1), trace);
        uint32_t **pc = malloc(sizeof(uint32_t *) * 2);
        uint32_t plogs[2] = {n, n};
        get_twiddles(n > 2 ? n : 3);
        for (int k = 0; k < 2; k++) {
            pc[k] = malloc(4 * N);
            memcpy(pc[k], trace + N * (W + k), 4 * N);
            circle_interpolate(pc[k], n);
        }
        commit_coeffs(&trees[0], 2, pc, plogs, B, &ch);
    } else {
        commit_coeffs(&trees[0], 0, NULL, NULL, B, &ch);
    }

    /* Tree 1: main trace: generate -> interpolate -> LDE -> Merkle (preparation.zig:113-210). */
    if (air == AIR_POSEIDON) ora_poseidon_trace(n, trace);
    else if (air == AIR_WIDEFIB) ora_widefib_trace(n, W, prm->trace_mode, prm->trace_seed, trace);
    double t1 = now_ms();
    uint32_t **tc = malloc(sizeof(uint32_t *) * W);
    uint32_t *tlogs = malloc(4 * W);
    get_twiddles(n > 2 ? n : 3);
    for (uint32_t c = 0; c < W; c++) {
        tc[c] = malloc(4 * N);
        memcpy(tc[c], trace + (uint64_t)c * N, 4 * N);
        tlogs[c] = n;
    }
    free(trace);
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t c = 0; c < (int64_t)W; c++) circle_interpolate(tc[c], n);
    commit_coeffs(&trees[1], W, tc, tlogs, B, &ch);
    free(tlogs);
    double t2 = now_ms();

    /* mixStatement (wide_fibonacci.zig:496-498; poseidon.zig:571-573: log_n_instances;
     * xor.zig:599-605: log_size, log_step, then mixU64(offset)). */
    uint32_t st[2] = {n, W};
    if (air == AIR_POSEIDON) st[0] = n + 3;
    if (air == AIR_XOR) st[1] = (uint32_t)xs[0];
    ch_mix_u32s(&ch, st, air == AIR_POSEIDON ? 1 : 2);
    if (air == AIR_XOR) ch_mix_u64(&ch, xs[1]);

    /* prove.zig:209-318: composition. */
    qm31 alpha = ch_draw_secure_felt(&ch);
    const uint32_t NC = air != AIR_WIDEFIB ? 1 : W - 2, CL = n + 1; /* composition log size */
    uint64_t CN = 1ull << CL;
    qm31 *pw = malloc(sizeof(qm31) * NC);
    pw[0] = Q_ONE;
    for (uint32_t k = 1; k < NC; k++) pw[k] = q_mul(pw[k - 1], alpha);
    /* Trace evaluations on CanonicCoset(n+1): the LDE itself when blowup is 1. */
    uint32_t **ev = malloc(sizeof(uint32_t *) * W);
    int own_ev = (B != 1) && air == AIR_WIDEFIB;
    for (uint32_t c = 0; c < W; c++) {
        if (!own_ev) ev[c] = trees[1].lde[c];
        else ev[c] = malloc(4 * CN);
    }
    if (own_ev) {
        get_twiddles(CL);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t c = 0; c < (int64_t)W; c++) ora_circle_lde(trees[1].coeffs[c], n, 1, ev[c]);
    }
    /* cosetVanishing(CanonicCoset(n).coset(), domain.at(0/1)) (constraints.zig:16-27). */
    uint32_t den_inv[2];
    for (int r = 0; r < 2; r++) {
        coset tc_ = coset_odds(n);
        pt dp = domain_at(CL, (uint64_t)r);
        pt shifted = p_add(p_add(dp, p_conj(idx_to_point(tc_.initial_index))),
                           idx_to_point(tc_.step_size / 2)); /* half_step */
        uint32_t x = shifted.x;
        for (uint32_t i = 1; i < n; i++) x = m_double_x(x);
        den_inv[r] = m_inv(x);
    }
    uint32_t *comp[4];
    for (int k = 0; k < 4; k++) comp[k] = malloc(4 * CN);
    if (air != AIR_WIDEFIB) {
        /* accumulateColumn of a constant column (accumulation.zig:142-159): the one
         * constraint's coefficient is pw[0] = 1. */
        uint32_t cv[4] = {n + 3, POS_COLS_PER_REP, POS_COLS, 1};
        if (air == AIR_XOR) {
            cv[0] = n;
            cv[1] = (uint32_t)xs[0];
            cv[2] = (uint32_t)(xs[1] % P);
        }
        for (int k = 0; k < 4; k++)
            for (uint64_t row = 0; row < CN; row++) comp[k][row] = cv[k];
    } else {
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < (int64_t)CN; row++) {
        uint32_t a = ev[0][row], b = ev[1][row];
        qm31 acc = Q_ZERO;
        for (uint32_t i = 0; i < NC; i++) {
            uint32_t c = ev[i + 2][row];
            uint32_t cons = m_sub(c, m_add(m_mul(a, a), m_mul(b, b)));
            acc = q_add(acc, q_mul_m(pw[NC - 1 - i], cons));
            a = b;
            b = c;
        }
        acc = q_mul_m(acc, den_inv[(uint64_t)row >> n]);
        comp[0][row] = acc.c0.a; comp[1][row] = acc.c0.b; comp[2][row] = acc.c1.a; comp[3][row] = acc.c1.b;
    }
    }
    if (own_ev)
        for (uint32_t c = 0; c < W; c++) free(ev[c]);
    free(ev);
    free(pw);
    /* interpolate + splitAtMid (secure_poly.zig:164-215; poly.zig:302-318) -> commitPolys. */
    get_twiddles(CL);
#pragma omp parallel for
    for (int k = 0; k < 4; k++) circle_interpolate(comp[k], CL);
    uint32_t **cc = malloc(sizeof(uint32_t *) * 8);
    uint32_t clogs[8];
    for (int k = 0; k < 4; k++) {
        cc[k] = malloc(4 * N);
        cc[4 + k] = malloc(4 * N);
        memcpy(cc[k], comp[k], 4 * N);
        memcpy(cc[4 + k], comp[k] + N, 4 * N);
        clogs[k] = clogs[4 + k] = n;
        free(comp[k]);
    }
    commit_coeffs(&trees[2], 8, cc, clogs, B, &ch);
    double t3 = now_ms();

    /* OODS point (circle.zig:159-167); every column is sampled once at it. */
    qpt oods = draw_oods_point(&ch);
    uint32_t total_cols = trees[0].n_cols + trees[1].n_cols + trees[2].n_cols;
    uint32_t *ns = malloc(4 * total_cols);
    qpt *pts = malloc(sizeof(qpt) * 2 * total_cols);
    for (uint32_t i = 0; i < total_cols; i++) {
        ns[i] = 1;
        pts[2 * i] = oods;
    }
    double tm[3] = {t0, t1, t2};
    int rc = prove_tail(&ch, trees, 3, ns, pts, prm, tm, t3, out, cap, out_len);
    free(ns);
    free(pts);
    for (int tr = 0; tr < 3; tr++) ctree_free(&trees[tr]);
    return rc;
}

int ora_prove_wide_fibonacci(const ora_widefib_params *prm, uint8_t *out, size_t cap, size_t *out_len) {
    return prove_core(AIR_WIDEFIB, prm, NULL, out, cap, out_len);
}

/* xor.prove (src/examples/xor.zig) for the statement (log_size, log_step, offset); prm
 * supplies the PCS configuration only. */
int ora_prove_xor(uint32_t log_size, uint32_t log_step, uint64_t offset, const ora_widefib_params *prm,
                  uint8_t *out, size_t cap, size_t *out_len) {
    if (log_size < 1 || log_size > 30 || log_step > log_size) return 1;
    ora_widefib_params p = *prm;
    p.log_n_rows = log_size;
    p.sequence_len = 1;
    const uint64_t xs[2] = {log_step, offset};
    return prove_core(AIR_XOR, &p, xs, out, cap, out_len);
}

/* poseidon.prove (src/examples/poseidon.zig:67-73): 2^(log_n_instances - 3) rows of 1264
 * columns; the trace fields of prm are ignored. */
int ora_prove_poseidon(uint32_t log_n_instances, const ora_widefib_params *prm, uint8_t *out, size_t cap,
                       size_t *out_len) {
    if (log_n_instances < 4 || log_n_instances > 33) return 1;
    ora_widefib_params p = *prm;
    p.log_n_rows = log_n_instances - 3;
    p.sequence_len = POS_COLS;
    return prove_core(AIR_POSEIDON, &p, NULL, out, cap, out_len);
}

/* ======================================================================
 * LogUp / GKR (src/core/fraction.zig:9-73; src/frontends/riscv/air/logup.zig:74-105;
 * src/prover/lookups/{gkr_circuit,gkr_prover,sumcheck,mle,utils}.zig) and the
 * XOR-lookup AIR of the LogUp workload (defined by this backend's LogUp benchmark: a, b uniform
 * 8-bit from splitmix64, c = a ^ b; table of all 2^16 (x, y, x ^ y) with
 * multiplicities; denominators z - (a + b*alpha + c*alpha^2)).
 * ==================================================================== */
static inline uint64_t ora_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

void ora_xor_trace(uint32_t log_size, uint64_t seed, uint32_t *a, uint32_t *b, uint32_t *c, uint32_t *mult) {
    uint64_t n = 1ull << log_size;
    memset(mult, 0, sizeof(uint32_t) << 16);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t w = ora_splitmix64(seed ^ (i * 0xD1B54A32D192ED03ull));
        a[i] = (uint32_t)(w & 0xff);
        b[i] = (uint32_t)((w >> 8) & 0xff);
        c[i] = a[i] ^ b[i];
        mult[(a[i] << 8) | b[i]]++;
    }
}

static inline qm31 logup_combine(uint32_t a, uint32_t b, uint32_t c, qm31 alpha, qm31 alpha2) {
    return q_add(q_add(q_from_m(a), q_mul_m(alpha, b)), q_mul_m(alpha2, c));
}

/* den[i] = z - (a[i] + b[i]*alpha + c[i]*alpha^2); table rows when a == NULL. */
void ora_logup_denominators(const uint32_t *a, const uint32_t *b, const uint32_t *c, uint64_t n, const uint32_t z[4],
                            const uint32_t alpha[4], uint32_t *den) {
    qm31 zz = q_load(z), al = q_load(alpha), al2 = q_square(al);
    for (uint64_t i = 0; i < n; i++) {
        uint32_t x, y, r;
        if (a) {
            x = a[i];
            y = b[i];
            r = c[i];
        } else {
            x = (uint32_t)((i >> 8) & 0xff);
            y = (uint32_t)(i & 0xff);
            r = x ^ y;
        }
        q_store(q_sub(zz, logup_combine(x, y, r, al, al2)), den + 4 * i);
    }
}

/* Cumulative column in trace order (logup.zig:74-91): S_i = sum_{j<=i} num_j / den_j.
 * num: QM31 AoS (num_kind 0), M31 (num_kind 1) or all ones (num_kind 2).
 * Returns -1 on a zero denominator. */
int ora_logup_cumulative(int num_kind, const uint32_t *num, const uint32_t *den, uint64_t n, uint32_t *sums,
                         uint32_t claimed[4]) {
    qm31 acc = Q_ZERO;
    for (uint64_t i = 0; i < n; i++) {
        qm31 d = q_load(den + 4 * i);
        if (q_is_zero(d)) return -1;
        qm31 nu = num_kind == 0 ? q_load(num + 4 * i) : num_kind == 1 ? q_from_m(num[i]) : Q_ONE;
        acc = q_add(acc, q_mul(nu, q_inv(d)));
        if (sums) q_store(acc, sums + 4 * i);
    }
    q_store(acc, claimed);
    return 0;
}

/* ======================================================================
 * XOR-lookup AIR as a whole STARK proof (defined by this backend; the LogUp machinery follows
 * src/frontends/riscv/air/logup.zig:74-105 and the lookup-table component of
 * src/frontends/riscv/air/lookups/tables/component.zig):
 *   tree 0 (preprocessed): is_first, the table tuple (tx, ty, tz) = ((i >> 8) & 255, i & 255, xor)
 *   tree 1 (main): a, b, c = a ^ b (uniform 8-bit lookups) and the table multiplicities m
 *   tree 2 (interaction): the cumulative column S of the row pairs (1 / d1) + (-m / d2),
 *           d1 = z - (a + alpha b + alpha^2 c), d2 = z - (tx + alpha ty + alpha^2 tz)
 *   tree 3: the composition of the pair constraint (logup.zig:93-105)
 *           (S(x) - S(x - g) + is_first(x) * claimed) * d1 * d2 - (d2 - m * d1),
 *           g the step of the trace coset, so S(x - g) is the previous row.
 * Rows are in coset order, row i stored at its bit-reversed circle-domain position; the
 * lookups balance the table, so the claimed sum is zero (the statement).  Transcript: config,
 * root 0, root 1, mixU32s([log_n]), z, alpha, root 2, mixFelts([claimed]), composition
 * coefficient, root 3, OODS point; masks: trees 0, 1 and 3 at the OODS point, the interaction
 * columns at the OODS point and at it minus g.
 * ==================================================================== */
static uint64_t prev_row_position(uint32_t lg, uint64_t r) {
    /* position of x - g (g = CanonicCoset(lg - 1).step) for the point x at bit-reversed
     * position r of CanonicCoset(lg).circleDomain() */
    uint64_t half = 1ull << (lg - 1), i = bitrev(r, lg), j;
    if (i < half) j = (i + half - 1) & (half - 1);
    else j = half + ((i - half + 1) & (half - 1));
    return bitrev(j, lg);
}

int ora_prove_logup_xor(uint32_t log_n, uint64_t seed, const ora_widefib_params *prm, uint8_t *out, size_t cap,
                        size_t *out_len) {
    const uint32_t n = log_n, B = prm->log_blowup_factor;
    if (n < 16 || n > 26 || B < 1) return 1;
    double t0 = now_ms();
    memset(g_timings, 0, sizeof g_timings);
    init_gen_doubles();
    channel ch;
    memset(&ch, 0, sizeof ch);
    qm31 cfg = {{prm->pow_bits, B}, {prm->n_queries, prm->log_last_layer_degree_bound}};
    ch_mix_felts(&ch, &cfg, 1);
    const uint64_t N = 1ull << n;
    uint32_t *ar = malloc(4 * N), *br = malloc(4 * N), *cr = malloc(4 * N), *mult = malloc(4u << 16);
    ora_xor_trace(n, seed, ar, br, cr, mult);
    uint32_t **pc = malloc(sizeof(uint32_t *) * 4), **mc = malloc(sizeof(uint32_t *) * 4);
    for (int k = 0; k < 4; k++) {
        pc[k] = malloc(4 * N);
        mc[k] = malloc(4 * N);
    }
    for (uint64_t i = 0; i < N; i++) {
        uint64_t p = circle_bitrev_index(n, i);
        uint32_t tx = (uint32_t)((i >> 8) & 0xff), ty = (uint32_t)(i & 0xff);
        pc[0][p] = i == 0;
        pc[1][p] = tx;
        pc[2][p] = ty;
        pc[3][p] = tx ^ ty;
        mc[0][p] = ar[i];
        mc[1][p] = br[i];
        mc[2][p] = cr[i];
        mc[3][p] = i < (1u << 16) ? mult[i] : 0;
    }
    double t1 = now_ms();
    ctree trees[4];
    memset(trees, 0, sizeof trees);
    uint32_t logs4[4] = {n, n, n, n};
    get_twiddles(n + B > n + 1 ? n + B : n + 1);
#pragma omp parallel for
    for (int k = 0; k < 8; k++) circle_interpolate(k < 4 ? pc[k] : mc[k - 4], n);
    commit_coeffs(&trees[0], 4, pc, logs4, B, &ch);
    commit_coeffs(&trees[1], 4, mc, logs4, B, &ch);
    double t2 = now_ms();
    ch_mix_u32s(&ch, &n, 1);
    qm31 z = ch_draw_secure_felt(&ch), al = ch_draw_secure_felt(&ch), al2 = q_square(al);
    /* the interaction column in coset order */
    qm31 *d1 = malloc(sizeof(qm31) * N), *d2 = malloc(sizeof(qm31) * N);
    uint32_t **sc = malloc(sizeof(uint32_t *) * 4);
    for (int k = 0; k < 4; k++) sc[k] = malloc(4 * N);
    qm31 acc = Q_ZERO;
    int zero_den = 0;
    for (uint64_t i = 0; i < N; i++) {
        uint32_t tx = (uint32_t)((i >> 8) & 0xff), ty = (uint32_t)(i & 0xff);
        d1[i] = q_sub(z, logup_combine(ar[i], br[i], cr[i], al, al2));
        d2[i] = q_sub(z, logup_combine(tx, ty, tx ^ ty, al, al2));
        qm31 den = q_mul(d1[i], d2[i]);
        if (q_is_zero(den)) zero_den = 1;
        uint32_t m = i < (1u << 16) ? mult[i] : 0;
        qm31 num = q_sub(d2[i], q_mul_m(d1[i], m));
        if (!zero_den) acc = q_add(acc, q_mul(num, q_inv(den)));
        uint64_t p = circle_bitrev_index(n, i);
        sc[0][p] = acc.c0.a;
        sc[1][p] = acc.c0.b;
        sc[2][p] = acc.c1.a;
        sc[3][p] = acc.c1.b;
    }
    free(d1);
    free(d2);
    free(ar);
    free(br);
    free(cr);
    free(mult);
    const qm31 claimed = acc;
    if (zero_den || !q_is_zero(claimed)) {
        for (int k = 0; k < 4; k++) free(sc[k]);
        free(sc);
        for (int tr = 0; tr < 2; tr++) ctree_free(&trees[tr]);
        return 5; /* the lookups do not balance the table */
    }
#pragma omp parallel for
    for (int k = 0; k < 4; k++) circle_interpolate(sc[k], n);
    commit_coeffs(&trees[2], 4, sc, logs4, B, &ch);
    ch_mix_felts(&ch, &claimed, 1);
    (void)ch_draw_secure_felt(&ch); /* composition coefficient: one constraint, power 1 */

    /* composition on CanonicCoset(n + 1) */
    const uint32_t CL = n + 1;
    const uint64_t CN = 1ull << CL;
    const uint32_t *ev[12];
    uint32_t *own[12] = {0};
    for (int k = 0; k < 12; k++) {
        const ctree *t = &trees[k / 4];
        if (B == 1) {
            ev[k] = t->lde[k % 4];
        } else {
            own[k] = malloc(4 * CN);
            ora_circle_lde(t->coeffs[k % 4], n, 1, own[k]);
            ev[k] = own[k];
        }
    }
    uint32_t den_inv[2];
    for (int r = 0; r < 2; r++) {
        coset tc_ = coset_odds(n);
        pt dp = domain_at(CL, (uint64_t)r);
        pt shifted = p_add(p_add(dp, p_conj(idx_to_point(tc_.initial_index))), idx_to_point(tc_.step_size / 2));
        uint32_t x = shifted.x;
        for (uint32_t i = 1; i < n; i++) x = m_double_x(x);
        den_inv[r] = m_inv(x);
    }
    uint32_t *comp[4];
    for (int k = 0; k < 4; k++) comp[k] = malloc(4 * CN);
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < (int64_t)CN; r++) {
        uint64_t q = prev_row_position(CL, (uint64_t)r);
        qm31 S = {{ev[8][r], ev[9][r]}, {ev[10][r], ev[11][r]}};
        qm31 Sp = {{ev[8][q], ev[9][q]}, {ev[10][q], ev[11][q]}};
        qm31 e1 = q_sub(z, logup_combine(ev[4][r], ev[5][r], ev[6][r], al, al2));
        qm31 e2 = q_sub(z, logup_combine(ev[1][r], ev[2][r], ev[3][r], al, al2));
        qm31 delta = q_add(q_sub(S, Sp), q_mul_m(claimed, ev[0][r]));
        qm31 cons = q_sub(q_mul(q_mul(delta, e1), e2), q_sub(e2, q_mul_m(e1, ev[7][r])));
        cons = q_mul_m(cons, den_inv[(uint64_t)r >> n]);
        comp[0][r] = cons.c0.a;
        comp[1][r] = cons.c0.b;
        comp[2][r] = cons.c1.a;
        comp[3][r] = cons.c1.b;
    }
    for (int k = 0; k < 12; k++) free(own[k]);
    get_twiddles(CL);
#pragma omp parallel for
    for (int k = 0; k < 4; k++) circle_interpolate(comp[k], CL);
    uint32_t **cc = malloc(sizeof(uint32_t *) * 8);
    uint32_t clogs[8];
    for (int k = 0; k < 4; k++) {
        cc[k] = malloc(4 * N);
        cc[4 + k] = malloc(4 * N);
        memcpy(cc[k], comp[k], 4 * N);
        memcpy(cc[4 + k], comp[k] + N, 4 * N);
        clogs[k] = clogs[4 + k] = n;
        free(comp[k]);
    }
    commit_coeffs(&trees[3], 8, cc, clogs, B, &ch);
    double t3 = now_ms();

    qpt oods = draw_oods_point(&ch);
    qpt prev = qp_add_m(oods, idx_to_point(idx_neg(subgroup_gen(n))));
    uint32_t ns[20];
    qpt pts[40];
    for (int i = 0; i < 20; i++) {
        ns[i] = (i >= 8 && i < 12) ? 2 : 1;
        pts[2 * i] = oods;
        pts[2 * i + 1] = prev;
    }
    double tm[3] = {t0, t1, t2};
    int rc = prove_tail(&ch, trees, 4, ns, pts, prm, tm, t3, out, cap, out_len);
    for (int tr = 0; tr < 4; tr++) ctree_free(&trees[tr]);
    return rc;
}

/* EqEvals over the hypercube: out[i] = scale * prod_j (bit_{n-1-j}(i) ? y_j : 1 - y_j)
 * (gkr_circuit.zig:651-674: y[0] selects the upper half). */
void ora_gkr_gen_eq_evals(const uint32_t *y, uint32_t n, const uint32_t scale[4], uint32_t *out) {
    uint64_t len = 1ull << n;
    q_store(q_load(scale), out);
    uint64_t cur = 1;
    for (int32_t j = (int32_t)n - 1; j >= 0; j--) {
        qm31 yj = q_load(y + 4 * j), y0 = q_sub(Q_ONE, yj);
        for (uint64_t i = 0; i < cur; i++) {
            qm31 v = q_load(out + 4 * i);
            q_store(q_mul(v, yj), out + 4 * (i + cur));
            q_store(q_mul(v, y0), out + 4 * i);
        }
        cur *= 2;
    }
    (void)len;
}

/* Layer kinds (gkr_circuit.zig:79-92). */
enum { GKR_GRAND_PRODUCT = 0, GKR_LOGUP_GENERIC = 1, GKR_LOGUP_MULTIPLICITIES = 2, GKR_LOGUP_SINGLES = 3 };

static inline qm31 layer_num(int kind, const uint32_t *num, uint64_t i) {
    if (kind == GKR_LOGUP_GENERIC) return q_load(num + 4 * i);
    if (kind == GKR_LOGUP_MULTIPLICITIES) return q_from_m(num[i]);
    return Q_ONE;
}

/* nextLayer (gkr_circuit.zig:449-503): pairs (2i, 2i+1); LogUp layers become generic. */
void ora_gkr_next_layer(int kind, const uint32_t *num, const uint32_t *den, uint32_t n_vars, uint32_t *out_num,
                        uint32_t *out_den) {
    uint64_t half = 1ull << (n_vars - 1);
    for (uint64_t i = 0; i < half; i++) {
        if (kind == GKR_GRAND_PRODUCT) {
            q_store(q_mul(q_load(den + 8 * i), q_load(den + 8 * i + 4)), out_den + 4 * i);
            continue;
        }
        qm31 n0 = layer_num(kind, num, 2 * i), n1 = layer_num(kind, num, 2 * i + 1);
        qm31 d0 = q_load(den + 8 * i), d1 = q_load(den + 8 * i + 4);
        q_store(q_add(q_mul(d1, n0), q_mul(d0, n1)), out_num + 4 * i);
        q_store(q_mul(d0, d1), out_den + 4 * i);
    }
}

/* Mle.fixFirstVariable (mle.zig:80-95; utils.zig:234-244): out[i] = a*(e[i+h]-e[i]) + e[i]. */
void ora_mle_fix_first(int base, const uint32_t *in, uint32_t n_vars, const uint32_t assignment[4], uint32_t *out) {
    uint64_t half = 1ull << (n_vars - 1);
    qm31 as = q_load(assignment);
    for (uint64_t i = 0; i < half; i++) {
        qm31 e0 = base ? q_from_m(in[i]) : q_load(in + 4 * i);
        qm31 e1 = base ? q_from_m(in[i + half]) : q_load(in + 4 * (i + half));
        q_store(q_add(q_mul(as, q_sub(e1, e0)), e0), out + 4 * i);
    }
}

/* --- univariate polynomials (utils.zig:14-214): coefficients ascending, trailing zeros trimmed --- */
typedef struct { qm31 c[8]; uint32_t len; } upoly;
static void up_trim(upoly *p) {
    while (p->len > 0 && q_is_zero(p->c[p->len - 1])) p->len--;
}
static qm31 up_eval(const upoly *p, qm31 x) {
    qm31 acc = Q_ZERO;
    for (uint32_t i = p->len; i > 0; i--) acc = q_add(q_mul(acc, x), p->c[i - 1]);
    return acc;
}
static qm31 q_div(qm31 a, qm31 b) { return q_mul(a, q_inv(b)); }
/* interpolateLagrange (utils.zig:147-190) for 4 points; the interpolant is unique. */
static upoly up_lagrange4(const qm31 xs[4], const qm31 ys[4]) {
    upoly acc = {.len = 0};
    for (int i = 0; i < 4; i++) {
        qm31 prod = ys[i];
        for (int j = 0; j < 4; j++)
            if (j != i) prod = q_div(prod, q_sub(xs[i], xs[j]));
        upoly term = {.len = 1};
        term.c[0] = prod;
        up_trim(&term);
        for (int j = 0; j < 4; j++) {
            if (j == i || term.len == 0) continue;
            upoly nt = {.len = term.len + 1};
            for (uint32_t k = 0; k < nt.len; k++) nt.c[k] = Q_ZERO;
            for (uint32_t k = 0; k < term.len; k++) {
                nt.c[k] = q_add(nt.c[k], q_mul(term.c[k], q_neg(xs[j])));
                nt.c[k + 1] = q_add(nt.c[k + 1], term.c[k]);
            }
            up_trim(&nt);
            term = nt;
        }
        uint32_t n = acc.len > term.len ? acc.len : term.len;
        upoly s = {.len = n};
        for (uint32_t k = 0; k < n; k++)
            s.c[k] = q_add(k < acc.len ? acc.c[k] : Q_ZERO, k < term.len ? term.c[k] : Q_ZERO);
        up_trim(&s);
        acc = s;
    }
    return acc;
}
static qm31 eq1(qm31 x, qm31 y) { /* utils.zig:220-231 */
    return q_add(q_mul(x, y), q_mul(q_sub(Q_ONE, x), q_sub(Q_ONE, y)));
}
/* correctSumAsPolyInFirstVariable (gkr_circuit.zig:390-429). */
static upoly gkr_correct_sum(qm31 f0, qm31 f2, qm31 claim, const qm31 *y, uint32_t ylen, uint32_t k) {
    qm31 eq_prefix = Q_ONE;
    for (uint32_t i = 0; i < ylen - k + 1; i++) eq_prefix = q_mul(eq_prefix, q_sub(Q_ONE, y[i]));
    qm31 a_const = q_div(Q_ONE, eq_prefix);
    qm31 yi = y[ylen - k];
    qm31 b_const = q_div(q_sub(Q_ONE, yi), q_sub(Q_ONE, q_add(yi, yi)));
    qm31 two = q_from_m(2);
    qm31 r0 = q_mul(q_mul(f0, q_sub(Q_ONE, yi)), a_const);
    qm31 r1 = q_sub(claim, r0);
    qm31 r2 = q_mul(q_mul(f2, eq1(two, yi)), a_const);
    qm31 xs[4] = {Q_ZERO, Q_ONE, two, b_const}, ys[4] = {r0, r1, r2, Q_ZERO};
    return up_lagrange4(xs, ys);
}

/* One GKR instance's sum-check oracle (gkr_circuit.zig:231-388). */
typedef struct {
    int kind;          /* current layer kind (multiplicities become generic once fixed) */
    uint32_t n_vars;   /* variables of the layer (oracle has n_vars - 1) */
    uint32_t *num;     /* M31 (multiplicities) or QM31 AoS */
    uint32_t *den;     /* QM31 AoS (values for grand product) */
    qm31 corr;         /* eq_fixed_var_correction */
} gkr_oracle;

/* sum over i < n_terms of eq[i] * term(i) at r = 0 and r = 2 (gkr_circuit.zig:505-643). */
static void gkr_eval_sums(const gkr_oracle *o, const qm31 *eq, qm31 lambda, qm31 out[2]) {
    uint64_t nt = 1ull << (o->n_vars - 2);
    qm31 e0 = Q_ZERO, e2 = Q_ZERO;
    for (uint64_t i = 0; i < nt; i++) {
        qm31 t0, t2;
        if (o->kind == GKR_GRAND_PRODUCT) {
            qm31 a0 = q_load(o->den + 8 * i), a1 = q_load(o->den + 8 * i + 4);
            qm31 b0 = q_load(o->den + 8 * (nt + i)), b1 = q_load(o->den + 8 * (nt + i) + 4);
            qm31 c0 = q_sub(q_add(b0, b0), a0), c1 = q_sub(q_add(b1, b1), a1);
            t0 = q_mul(a0, a1);
            t2 = q_mul(c0, c1);
        } else {
            qm31 d00 = q_load(o->den + 8 * i), d01 = q_load(o->den + 8 * i + 4);
            qm31 d10 = q_load(o->den + 8 * (nt + i)), d11 = q_load(o->den + 8 * (nt + i) + 4);
            qm31 d20 = q_sub(q_add(d10, d10), d00), d21 = q_sub(q_add(d11, d11), d01);
            qm31 n00 = layer_num(o->kind, o->num, 2 * i), n01 = layer_num(o->kind, o->num, 2 * i + 1);
            qm31 n10 = layer_num(o->kind, o->num, 2 * (nt + i)), n11 = layer_num(o->kind, o->num, 2 * (nt + i) + 1);
            qm31 n20 = q_sub(q_add(n10, n10), n00), n21 = q_sub(q_add(n11, n11), n01);
            qm31 fn0 = q_add(q_mul(d01, n00), q_mul(d00, n01)), fd0 = q_mul(d00, d01);
            qm31 fn2 = q_add(q_mul(d21, n20), q_mul(d20, n21)), fd2 = q_mul(d20, d21);
            t0 = q_add(fn0, q_mul(lambda, fd0));
            t2 = q_add(fn2, q_mul(lambda, fd2));
        }
        e0 = q_add(e0, q_mul(eq[i], t0));
        e2 = q_add(e2, q_mul(eq[i], t2));
    }
    out[0] = e0;
    out[1] = e2;
}

static void gkr_fix(gkr_oracle *o, qm31 r) {
    uint32_t nv = o->n_vars;
    uint64_t half = 1ull << (nv - 1);
    uint32_t *nd = malloc(16 * half);
    ora_mle_fix_first(0, o->den, nv, (uint32_t[4]){r.c0.a, r.c0.b, r.c1.a, r.c1.b}, nd);
    free(o->den);
    o->den = nd;
    if (o->kind == GKR_LOGUP_GENERIC || o->kind == GKR_LOGUP_MULTIPLICITIES) {
        uint32_t *nn = malloc(16 * half);
        ora_mle_fix_first(o->kind == GKR_LOGUP_MULTIPLICITIES, o->num, nv, (uint32_t[4]){r.c0.a, r.c0.b, r.c1.a, r.c1.b},
                          nn);
        free(o->num);
        o->num = nn;
        o->kind = GKR_LOGUP_GENERIC;
    }
    o->n_vars = nv - 1;
}

/* Word stream writer. */
typedef struct { uint32_t *w; size_t n, cap; } wbuf;
static void wb_u(wbuf *b, uint32_t v) {
    if (b->n < b->cap) b->w[b->n] = v;
    b->n++;
}
static void wb_q(wbuf *b, qm31 q) {
    wb_u(b, q.c0.a), wb_u(b, q.c0.b), wb_u(b, q.c1.a), wb_u(b, q.c1.b);
}

/* GKR batch prover (gkr_prover.zig:34-298, sumcheck.zig:76-197) over instances given as
 * input layers.  Returns the word count of the proof stream (written when it fits `cap`):
 *   per GKR layer: n_round_polys, each [len, coeffs]; n_masks, each [instance, n_cols, cols];
 *   then per instance: output claims [n, values]; ood point [n, values];
 *   per instance: claims to verify [n, values], n_vars; final channel digest (8 words). */
static size_t gkr_prove_on(channel *ch, uint32_t n_inst, const int *kinds, const uint32_t *const *nums,
                           const uint32_t *const *dens, const uint32_t *n_vars_in, uint32_t *out, size_t cap) {
    wbuf wb = {out, 0, cap};
    uint32_t n_layers = 0;
    for (uint32_t k = 0; k < n_inst; k++) n_layers = n_vars_in[k] > n_layers ? n_vars_in[k] : n_layers;
    /* layers[k][l] = layer with l variables of instance k (l = n_vars .. 0) */
    uint32_t ***lnum = calloc(n_inst, sizeof(uint32_t **)), ***lden = calloc(n_inst, sizeof(uint32_t **));
    int **lkind = calloc(n_inst, sizeof(int *));
    for (uint32_t k = 0; k < n_inst; k++) {
        uint32_t nv = n_vars_in[k];
        lnum[k] = calloc(nv + 1, sizeof(uint32_t *));
        lden[k] = calloc(nv + 1, sizeof(uint32_t *));
        lkind[k] = calloc(nv + 1, sizeof(int));
        uint64_t len = 1ull << nv;
        int kind = kinds[k];
        lkind[k][nv] = kind;
        lden[k][nv] = malloc(16 * len);
        memcpy(lden[k][nv], dens[k], 16 * len);
        if (kind == GKR_LOGUP_GENERIC) {
            lnum[k][nv] = malloc(16 * len);
            memcpy(lnum[k][nv], nums[k], 16 * len);
        } else if (kind == GKR_LOGUP_MULTIPLICITIES) {
            lnum[k][nv] = malloc(4 * len);
            memcpy(lnum[k][nv], nums[k], 4 * len);
        }
        for (int32_t l = (int32_t)nv - 1; l >= 0; l--) {
            uint64_t h = 1ull << l;
            lden[k][l] = malloc(16 * h);
            lnum[k][l] = kind == GKR_GRAND_PRODUCT ? NULL : malloc(16 * h);
            ora_gkr_next_layer(lkind[k][l + 1], lnum[k][l + 1], lden[k][l + 1], (uint32_t)l + 1, lnum[k][l],
                               lden[k][l]);
            lkind[k][l] = kind == GKR_GRAND_PRODUCT ? GKR_GRAND_PRODUCT : GKR_LOGUP_GENERIC;
        }
    }
    qm31 **claims = calloc(n_inst, sizeof(qm31 *)), **outputs = calloc(n_inst, sizeof(qm31 *));
    uint32_t *n_claims = calloc(n_inst, sizeof(uint32_t));
    qm31 *ood = NULL;
    uint32_t ood_len = 0;
    for (uint32_t layer = 0; layer < n_layers; layer++) {
        uint32_t remaining = n_layers - layer;
        for (uint32_t k = 0; k < n_inst; k++) {
            if (n_vars_in[k] != remaining) continue;
            /* output layer values (gkr_circuit.zig:162-190) */
            int gp = lkind[k][0] == GKR_GRAND_PRODUCT;
            n_claims[k] = gp ? 1 : 2;
            claims[k] = malloc(2 * sizeof(qm31));
            outputs[k] = malloc(2 * sizeof(qm31));
            if (gp) {
                claims[k][0] = q_load(lden[k][0]);
            } else {
                claims[k][0] = q_load(lnum[k][0]);
                claims[k][1] = q_load(lden[k][0]);
            }
            memcpy(outputs[k], claims[k], n_claims[k] * sizeof(qm31));
        }
        for (uint32_t k = 0; k < n_inst; k++)
            if (claims[k]) ch_mix_felts(ch, claims[k], n_claims[k]);
        /* EqEvals.generate(ood) = genEqEvals(ood[1..], 1 - ood[0]) */
        uint64_t eq_len = ood_len ? (1ull << (ood_len - 1)) : 1;
        qm31 *eq = malloc(sizeof(qm31) * eq_len);
        if (ood_len == 0) {
            eq[0] = Q_ONE;
        } else {
            uint32_t *yw = malloc(16 * ood_len);
            for (uint32_t i = 0; i < ood_len; i++) q_store(ood[i], yw + 4 * i);
            qm31 sc = q_sub(Q_ONE, ood[0]);
            uint32_t scw[4];
            q_store(sc, scw);
            uint32_t *ew = malloc(16 * eq_len);
            ora_gkr_gen_eq_evals(yw + 4, ood_len - 1, scw, ew);
            for (uint64_t i = 0; i < eq_len; i++) eq[i] = q_load(ew + 4 * i);
            free(ew);
            free(yw);
        }
        qm31 sc_alpha = ch_draw_secure_felt(ch);
        qm31 lambda = ch_draw_secure_felt(ch);
        /* sum-check instances: those with claims, layer with `layer + 1` variables */
        uint32_t n_sc = 0;
        uint32_t *sc_inst = malloc(sizeof(uint32_t) * n_inst);
        gkr_oracle *orc = calloc(n_inst, sizeof(gkr_oracle));
        qm31 *sc_claims = malloc(sizeof(qm31) * n_inst);
        for (uint32_t k = 0; k < n_inst; k++) {
            if (!claims[k]) continue;
            uint32_t lv = n_vars_in[k] - remaining + 1; /* layer of this instance with lv variables */
            gkr_oracle o;
            o.kind = lkind[k][lv];
            o.n_vars = lv;
            uint64_t len = 1ull << lv;
            o.den = malloc(16 * len);
            memcpy(o.den, lden[k][lv], 16 * len);
            o.num = NULL;
            if (o.kind == GKR_LOGUP_GENERIC) {
                o.num = malloc(16 * len);
                memcpy(o.num, lnum[k][lv], 16 * len);
            } else if (o.kind == GKR_LOGUP_MULTIPLICITIES) {
                o.num = malloc(4 * len);
                memcpy(o.num, lnum[k][lv], 4 * len);
            }
            o.corr = Q_ONE;
            /* randomLinearCombination(claims, lambda) = Horner */
            qm31 cl = Q_ZERO;
            for (uint32_t i = n_claims[k]; i > 0; i--) cl = q_add(q_mul(cl, lambda), claims[k][i - 1]);
            orc[n_sc] = o;
            sc_claims[n_sc] = cl;
            sc_inst[n_sc++] = k;
        }
        /* sumcheck.proveBatch */
        uint32_t sc_nv = 0;
        for (uint32_t i = 0; i < n_sc; i++) sc_nv = orc[i].n_vars - 1 > sc_nv ? orc[i].n_vars - 1 : sc_nv;
        for (uint32_t i = 0; i < n_sc; i++) {
            uint32_t unused = sc_nv - (orc[i].n_vars - 1);
            sc_claims[i] = q_mul_m(sc_claims[i], m_pow(2, unused));
        }
        wb_u(&wb, sc_nv);
        qm31 *assign = malloc(sizeof(qm31) * (sc_nv + 1));
        upoly *rp = malloc(sizeof(upoly) * (n_sc ? n_sc : 1));
        for (uint32_t round = 0; round < sc_nv; round++) {
            uint32_t rem = sc_nv - round;
            for (uint32_t i = 0; i < n_sc; i++) {
                uint32_t onv = orc[i].n_vars - 1;
                if (rem == onv) {
                    qm31 ev[2];
                    gkr_eval_sums(&orc[i], eq, lambda, ev);
                    ev[0] = q_mul(ev[0], orc[i].corr);
                    ev[1] = q_mul(ev[1], orc[i].corr);
                    rp[i] = gkr_correct_sum(ev[0], ev[1], sc_claims[i], ood, ood_len, onv);
                } else {
                    rp[i].len = 1;
                    rp[i].c[0] = q_mul_m(sc_claims[i], m_inv(2));
                    up_trim(&rp[i]);
                }
            }
            /* combined round poly: Horner over polys with sc_alpha */
            upoly acc = {.len = 0};
            for (uint32_t i = n_sc; i > 0; i--) {
                upoly sc = acc;
                for (uint32_t k = 0; k < sc.len; k++) sc.c[k] = q_mul(sc.c[k], sc_alpha);
                if (q_is_zero(sc_alpha)) sc.len = 0;
                up_trim(&sc);
                uint32_t n = sc.len > rp[i - 1].len ? sc.len : rp[i - 1].len;
                upoly s = {.len = n};
                for (uint32_t k = 0; k < n; k++)
                    s.c[k] = q_add(k < sc.len ? sc.c[k] : Q_ZERO, k < rp[i - 1].len ? rp[i - 1].c[k] : Q_ZERO);
                up_trim(&s);
                acc = s;
            }
            ch_mix_felts(ch, acc.c, acc.len);
            qm31 r = ch_draw_secure_felt(ch);
            wb_u(&wb, acc.len);
            for (uint32_t k = 0; k < acc.len; k++) wb_q(&wb, acc.c[k]);
            for (uint32_t i = 0; i < n_sc; i++) sc_claims[i] = up_eval(&rp[i], r);
            for (uint32_t i = 0; i < n_sc; i++) {
                if (rem != orc[i].n_vars - 1) continue;
                uint32_t onv = orc[i].n_vars - 1;
                qm31 z0 = ood[ood_len - onv];
                orc[i].corr = q_mul(orc[i].corr, eq1(r, z0));
                gkr_fix(&orc[i], r);
            }
            assign[round] = r;
        }
        /* masks (tryIntoMask) and their transcript */
        wb_u(&wb, n_sc);
        qm31 (*masks)[2][2] = malloc(sizeof(qm31[2][2]) * (n_sc ? n_sc : 1));
        uint32_t *mcols = malloc(sizeof(uint32_t) * (n_sc ? n_sc : 1));
        for (uint32_t i = 0; i < n_sc; i++) {
            gkr_oracle *o = &orc[i];
            if (o->kind == GKR_GRAND_PRODUCT) {
                mcols[i] = 1;
                masks[i][0][0] = q_load(o->den);
                masks[i][0][1] = q_load(o->den + 4);
            } else {
                mcols[i] = 2;
                masks[i][0][0] = layer_num(o->kind, o->num, 0);
                masks[i][0][1] = layer_num(o->kind, o->num, 1);
                masks[i][1][0] = q_load(o->den);
                masks[i][1][1] = q_load(o->den + 4);
            }
            qm31 flat[4];
            for (uint32_t c2 = 0; c2 < mcols[i]; c2++) {
                flat[2 * c2] = masks[i][c2][0];
                flat[2 * c2 + 1] = masks[i][c2][1];
            }
            ch_mix_felts(ch, flat, 2 * mcols[i]);
            wb_u(&wb, sc_inst[i]);
            wb_u(&wb, mcols[i]);
            for (uint32_t c2 = 0; c2 < 2 * mcols[i]; c2++) wb_q(&wb, flat[c2]);
        }
        qm31 challenge = ch_draw_secure_felt(ch);
        qm31 *nood = malloc(sizeof(qm31) * (sc_nv + 1));
        memcpy(nood, assign, sizeof(qm31) * sc_nv);
        nood[sc_nv] = challenge;
        free(ood);
        ood = nood;
        ood_len = sc_nv + 1;
        for (uint32_t i = 0; i < n_sc; i++) {
            uint32_t k = sc_inst[i];
            n_claims[k] = mcols[i];
            for (uint32_t c2 = 0; c2 < mcols[i]; c2++)
                claims[k][c2] = q_add(q_mul(challenge, q_sub(masks[i][c2][1], masks[i][c2][0])), masks[i][c2][0]);
        }
        for (uint32_t i = 0; i < n_sc; i++) {
            free(orc[i].num);
            free(orc[i].den);
        }
        free(masks), free(mcols), free(assign), free(rp), free(orc), free(sc_claims), free(sc_inst), free(eq);
    }
    for (uint32_t k = 0; k < n_inst; k++) {
        uint32_t n = outputs[k] ? (lkind[k][0] == GKR_GRAND_PRODUCT ? 1 : 2) : 0;
        wb_u(&wb, n);
        for (uint32_t i = 0; i < n; i++) wb_q(&wb, outputs[k][i]);
    }
    wb_u(&wb, ood_len);
    for (uint32_t i = 0; i < ood_len; i++) wb_q(&wb, ood[i]);
    for (uint32_t k = 0; k < n_inst; k++) {
        wb_u(&wb, claims[k] ? n_claims[k] : 0);
        for (uint32_t i = 0; claims[k] && i < n_claims[k]; i++) wb_q(&wb, claims[k][i]);
        wb_u(&wb, n_vars_in[k]);
    }
    for (int i = 0; i < 8; i++) {
        uint32_t d;
        memcpy(&d, ch->digest + 4 * i, 4);
        wb_u(&wb, d);
    }
    for (uint32_t k = 0; k < n_inst; k++) {
        for (uint32_t l = 0; l <= n_vars_in[k]; l++) {
            free(lnum[k][l]);
            free(lden[k][l]);
        }
        free(lnum[k]), free(lden[k]), free(lkind[k]), free(claims[k]), free(outputs[k]);
    }
    free(lnum), free(lden), free(lkind), free(claims), free(outputs), free(n_claims), free(ood);
    return wb.n;
}

size_t ora_gkr_prove(uint32_t n_inst, const int *kinds, const uint32_t *const *nums, const uint32_t *const *dens,
                     const uint32_t *n_vars_in, const uint8_t digest_in[32], uint32_t *out, size_t cap) {
    channel ch;
    memcpy(ch.digest, digest_in, 32);
    ch.n_draws = 0;
    return gkr_prove_on(&ch, n_inst, kinds, nums, dens, n_vars_in, out, cap);
}

/* LogUp workload (XOR-lookup AIR) from trace to quotient commitment, the same sequence as the
 * device pipeline: trace + multiplicities (zero-padded to n rows) -> commit [a, b, c, m] ->
 * z, alpha -> denominators -> cumulative columns (claimed sums mixed) -> commit the 8 coordinate
 * columns -> batch GKR proof -> OODS point, sampled values of the 12 polynomials -> quotients at
 * the OODS point -> quotient commitment.  Output words: 3 roots, 2 claimed sums, 12 sampled
 * values, GKR word count and words, final channel digest.  Returns the word count. */
size_t ora_logup_xor_pipeline(uint32_t log_n, uint64_t seed, uint32_t *out, size_t cap) {
    const uint32_t n = log_n, B = 1, L = n + 1;
    const uint64_t N = 1ull << n;
    channel ch;
    memset(ch.digest, 0, 32);
    ch.n_draws = 0;
    {
        uint32_t st[3] = {n, (uint32_t)seed, (uint32_t)(seed >> 32)};
        ch_mix_u32s(&ch, st, 3);
    }
    uint32_t *a = malloc(4 * N), *b = malloc(4 * N), *c = malloc(4 * N), *mult = malloc(4u << 16);
    ora_xor_trace(n, seed, a, b, c, mult);
    uint32_t *mpad = calloc(N, 4), *negm = malloc(4 * N);
    for (uint64_t i = 0; i < N; i++) {
        mpad[i] = i < (1u << 16) ? mult[i] : 0;
        negm[i] = m_neg(mpad[i]);
    }
    /* main commitment */
    ctree tm, ti;
    {
        uint32_t **cf = malloc(sizeof(uint32_t *) * 4);
        const uint32_t *src[4] = {a, b, c, mpad};
        uint32_t logs[4] = {n, n, n, n};
        get_twiddles(n);
        for (int k = 0; k < 4; k++) {
            cf[k] = malloc(4 * N);
            memcpy(cf[k], src[k], 4 * N);
            circle_interpolate(cf[k], n);
        }
        commit_coeffs(&tm, 4, cf, logs, B, &ch);
    }
    qm31 z = ch_draw_secure_felt(&ch), alpha = ch_draw_secure_felt(&ch);
    uint32_t zw[4], aw[4];
    q_store(z, zw);
    q_store(alpha, aw);
    uint32_t *dl = malloc(16 * N), *dt = malloc(16 * N), *sl = malloc(16 * N), *st = malloc(16 * N);
    ora_logup_denominators(a, b, c, N, zw, aw, dl);
    ora_logup_denominators(NULL, NULL, NULL, N, zw, aw, dt);
    uint32_t clw[4], ctw[4];
    int bad = ora_logup_cumulative(2, NULL, dl, N, sl, clw) | ora_logup_cumulative(1, negm, dt, N, st, ctw);
    qm31 cs[2] = {q_load(clw), q_load(ctw)};
    if (bad || !q_is_zero(q_add(cs[0], cs[1]))) return 0;
    ch_mix_felts(&ch, cs, 2);
    {
        uint32_t **cf = malloc(sizeof(uint32_t *) * 8);
        uint32_t logs[8];
        for (int k = 0; k < 8; k++) {
            const uint32_t *s2 = k < 4 ? sl : st;
            cf[k] = malloc(4 * N);
            for (uint64_t i = 0; i < N; i++) cf[k][i] = s2[4 * i + (k & 3)];
            circle_interpolate(cf[k], n);
            logs[k] = n;
        }
        commit_coeffs(&ti, 8, cf, logs, B, &ch);
    }
    /* GKR */
    int kinds[2] = {GKR_LOGUP_SINGLES, GKR_LOGUP_MULTIPLICITIES};
    const uint32_t *nums[2] = {NULL, negm}, *dens[2] = {dl, dt};
    uint32_t nvs[2] = {n, n};
    channel ch_probe = ch;
    size_t gn = gkr_prove_on(&ch_probe, 2, kinds, nums, dens, nvs, NULL, 0);
    uint32_t *gw = malloc(4 * gn);
    gkr_prove_on(&ch, 2, kinds, nums, dens, nvs, gw, gn);
    /* OODS point and sampled values */
    qm31 tt = ch_draw_secure_felt(&ch);
    qm31 t2_ = q_square(tt);
    qm31 inv1 = q_inv(q_add(t2_, Q_ONE));
    qpt oods = {q_mul(q_sub(Q_ONE, t2_), inv1), q_mul(q_add(tt, tt), inv1)};
    qm31 sv[12];
    {
        qm31 *w = eval_weights(oods, n);
        for (int k = 0; k < 12; k++) sv[k] = dot_weights(k < 4 ? tm.coeffs[k] : ti.coeffs[k - 4], w, N);
        free(w);
    }
    ch_mix_felts(&ch, sv, 12);
    /* quotients at the OODS point */
    qm31 qa = ch_draw_secure_felt(&ch);
    qm31 lc[12], sum_a = Q_ZERO, sum_b = Q_ZERO, rp = Q_ONE;
    qm31 cc = q_sub(q_conj(oods.y), oods.y);
    for (int k = 0; k < 12; k++) {
        qm31 av = q_sub(q_conj(sv[k]), sv[k]);
        qm31 bv = q_sub(q_mul(sv[k], cc), q_mul(av, oods.y));
        lc[k] = q_mul(rp, cc);
        sum_a = q_add(sum_a, q_mul(rp, av));
        sum_b = q_add(sum_b, q_mul(rp, bv));
        rp = q_mul(rp, qa);
    }
    const uint64_t LN = 1ull << L;
    cm31 det = c_sub(c_mul(oods.x.c0, oods.y.c1), c_mul(oods.y.c0, oods.x.c1));
    qm31 *quot = malloc(sizeof(qm31) * LN);
#pragma omp parallel for schedule(static)
    for (int64_t row = 0; row < (int64_t)LN; row++) {
        pt dp = domain_at(L, bitrev((uint64_t)row, L));
        qm31 num = Q_ZERO;
        for (int k = 0; k < 12; k++) num = q_add(num, q_mul_m(lc[k], (k < 4 ? tm.lde[k] : ti.lde[k - 4])[row]));
        num = q_sub(num, q_add(q_mul_m(sum_a, dp.y), sum_b));
        cm31 den = c_add(c_sub(det, c_mul_m(oods.y.c1, dp.x)), c_mul_m(oods.x.c1, dp.y));
        quot[row] = q_mul_c(num, c_inv(den));
    }
    ora_tree *qt = commit_qm31_column(quot, L, NULL);
    uint8_t r3[32];
    ora_tree_root(qt, r3);
    ch_mix_root(&ch, r3);
    /* output */
    wbuf wb = {out, 0, cap};
    uint8_t r1[32], r2[32];
    ora_tree_root(tm.tree, r1);
    ora_tree_root(ti.tree, r2);
    const uint8_t *roots[3] = {r1, r2, r3};
    for (int r = 0; r < 3; r++)
        for (int i = 0; i < 8; i++) {
            uint32_t v;
            memcpy(&v, roots[r] + 4 * i, 4);
            wb_u(&wb, v);
        }
    wb_q(&wb, cs[0]);
    wb_q(&wb, cs[1]);
    for (int k = 0; k < 12; k++) wb_q(&wb, sv[k]);
    wb_u(&wb, (uint32_t)gn);
    for (size_t i = 0; i < gn; i++) wb_u(&wb, gw[i]);
    for (int i = 0; i < 8; i++) {
        uint32_t v;
        memcpy(&v, ch.digest + 4 * i, 4);
        wb_u(&wb, v);
    }
    ora_tree_free(qt);
    free(quot), free(gw), free(dl), free(dt), free(sl), free(st), free(a), free(b), free(c), free(mult), free(mpad),
        free(negm);
    ctree_free(&tm);
    ctree_free(&ti);
    return wb.n;
}

/* Batched primitive pipeline, the same sequence as the device graph: n_columns seeded
 * uniform-M31 evaluations of log sizes drawn from [log_min, log_max] -> interpolate -> LDE at
 * blowup -> one lifted Merkle commitment; channel <- main root (draw); then for each of n_folds
 * rounds: commit the QM31 line (4 coordinate columns), mix its root and draw alpha, fold_line;
 * the last line is committed and mixed (draw) too.  Output: main root, n_folds + 1 line roots,
 * final channel digest, draw counter.  Returns the word count. */
size_t ora_batched_pipeline(uint32_t n_columns, uint32_t log_min, uint32_t log_max, uint32_t log_blowup,
                            uint32_t fri_log, uint32_t n_folds, uint64_t seed, uint32_t *out, size_t cap) {
    uint32_t *logs = malloc(4 * n_columns), *lde_logs = malloc(4 * n_columns);
    uint32_t **lde = malloc(sizeof(uint32_t *) * n_columns);
    for (uint32_t i = 0; i < n_columns; i++) {
        logs[i] = log_min + (uint32_t)(ora_splitmix64(seed ^ 0xA5A5A5A5ull ^ (i * 0x9E3779B97F4A7C15ull)) %
                                       (log_max - log_min + 1));
        lde_logs[i] = logs[i] + log_blowup;
        get_twiddles(lde_logs[i] > 2 ? lde_logs[i] : 3);
    }
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < (int64_t)n_columns; i++) {
        uint64_t n = 1ull << logs[i];
        uint32_t *v = malloc(4 * n);
        for (uint64_t r = 0; r < n; r++) v[r] = (uint32_t)(ora_splitmix64(seed + ((uint64_t)i << 40) + r) % P);
        circle_interpolate(v, logs[i]);
        lde[i] = malloc(4ull << lde_logs[i]);
        ora_circle_lde(v, logs[i], log_blowup, lde[i]);
        free(v);
    }
    ora_tree *t = ora_merkle_commit((const uint32_t *const *)lde, lde_logs, n_columns);
    wbuf wb = {out, 0, cap};
    uint8_t root[32];
    ora_tree_root(t, root);
    channel ch;
    memset(ch.digest, 0, 32);
    ch.n_draws = 0;
    ch_mix_root(&ch, root);
    (void)ch_draw_secure_felt(&ch);
    for (int i = 0; i < 8; i++) {
        uint32_t v;
        memcpy(&v, root + 4 * i, 4);
        wb_u(&wb, v);
    }
    uint64_t n0 = 1ull << fri_log;
    qm31 *line = malloc(sizeof(qm31) * n0);
    for (uint64_t i = 0; i < 4 * n0; i++) ((uint32_t *)line)[i] = (uint32_t)(ora_splitmix64((seed ^ 0x51F15EEDull) + i) % P);
    for (uint32_t j = 0; j <= n_folds; j++) {
        uint32_t lg = fri_log - j;
        ora_tree *ft = commit_qm31_column(line, lg, NULL);
        uint8_t r[32];
        ora_tree_root(ft, r);
        ora_tree_free(ft);
        ch_mix_root(&ch, r);
        qm31 a = ch_draw_secure_felt(&ch);
        for (int i = 0; i < 8; i++) {
            uint32_t v;
            memcpy(&v, r + 4 * i, 4);
            wb_u(&wb, v);
        }
        if (j == n_folds) break;
        qm31 *next = malloc(sizeof(qm31) << (lg - 1));
        uint32_t aw[4];
        q_store(a, aw);
        ora_fold_line((const uint32_t *)line, lg, coset_half_odds(lg).initial_index, aw, (uint32_t *)next);
        free(line);
        line = next;
    }
    for (int i = 0; i < 8; i++) {
        uint32_t v;
        memcpy(&v, ch.digest + 4 * i, 4);
        wb_u(&wb, v);
    }
    wb_u(&wb, ch.n_draws);
    free(line);
    ora_tree_free(t);
    for (uint32_t i = 0; i < n_columns; i++) free(lde[i]);
    free(lde), free(logs), free(lde_logs);
    return wb.n;
}
