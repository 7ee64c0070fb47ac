// GKR batch prover with the whole transcript on the device (src/prover/lookups/
// gkr_prover.zig:34-298 with sumcheck.zig:76-197 and the oracle of
// gkr_circuit.zig:231-429).
//
// The layer stacks (nextLayer), the eq tables, the sum-check rounds, the
// channel (mixFelts / drawSecureFelt on one quad of lanes, chan.hpp) and the
// masks all run on the GPU; the host only enqueues launches and reads the
// transcript back once.  Per GKR layer L (L sum-check rounds):
//   * begin: newly active instances take their output claims, the claims are
//     mixed, alpha and lambda drawn, the batched claims formed; one lane per
//     round rho precomputes what the round polynomial needs from the eq point
//     y (the eq-prefix inverse prod_{i<=rho} 1/(1-y_i), the extra Lagrange node
//     b = (1-y_rho)/(1-2y_rho) and the basis coefficients for the nodes
//     0, 1, 2, b), so no round inverts anything;
//   * rounds: fixing the previous round's variable is fused into this round's
//     sums (in place: index j < h is written by the thread that reads j and
//     j + h); rounds on layers above 2^GD_SMALL_LOG values are one launch over
//     many workgroups whose last workgroup (atomic counter, no waiting) turns
//     the sums into the round polynomial, mixes it and draws the challenge;
//     smaller rounds, whole small layers and the layer end (masks, challenge,
//     next claims) run in one workgroup.
// The round polynomial r(x) through (0, r0), (1, claim - r0), (2, r2), (b, 0)
// (correctSumAsPolyInFirstVariable, gkr_circuit.zig:390-429) is unique, so the
// closed-form coefficients equal the reference's Lagrange interpolation.
//
// Proof word stream (what stwo_rocm_gkr_prove_batch returns):
//   per GKR layer: n_rounds, each [len, len QM31 coefficients];
//                  n_masks, each [instance, n_columns, 2*n_columns QM31];
//   per instance: [n, output claims]; [n, ood point];
//   per instance: [n, claims to verify], n_variables; final channel digest (8 words).
#include <memory>
#include <vector>

#include "chan.hpp"
#include "prover.hpp"

namespace sr {

namespace {

// Univariate polynomial with ascending QM31 coefficients, trailing zeros trimmed
// (utils.zig:14-214).
struct Poly {
    std::vector<qm31> c;
    void trim() {
        while (!c.empty() && qiszero(c.back())) c.pop_back();
    }
    qm31 eval(qm31 x) const {
        qm31 acc = qzero();
        for (size_t i = c.size(); i > 0; i--) acc = qadd(qmul(acc, x), c[i - 1]);
        return acc;
    }
};

qm31 qdiv(qm31 a, qm31 b) { return qmul(a, qinv(b)); }

// Lagrange interpolation through 4 points (utils.zig:147-190).
Poly lagrange(const qm31 xs[4], const qm31 ys[4]) {
    Poly acc;
    for (int i = 0; i < 4; i++) {
        qm31 scale = ys[i];
        for (int j = 0; j < 4; j++)
            if (j != i) scale = qdiv(scale, qsub(xs[i], xs[j]));
        Poly term;
        term.c = {scale};
        term.trim();
        for (int j = 0; j < 4 && !term.c.empty(); j++) {
            if (j == i) continue;
            Poly nt;
            nt.c.assign(term.c.size() + 1, qzero());
            for (size_t k = 0; k < term.c.size(); k++) {
                nt.c[k] = qadd(nt.c[k], qmul(term.c[k], qneg(xs[j])));
                nt.c[k + 1] = qadd(nt.c[k + 1], term.c[k]);
            }
            nt.trim();
            term = nt;
        }
        Poly s;
        s.c.assign(std::max(acc.c.size(), term.c.size()), qzero());
        for (size_t k = 0; k < s.c.size(); k++)
            s.c[k] = qadd(k < acc.c.size() ? acc.c[k] : qzero(), k < term.c.size() ? term.c[k] : qzero());
        s.trim();
        acc = s;
    }
    return acc;
}

qm31 eq_one(qm31 x, qm31 y) { return qadd(qmul(x, y), qmul(qsub(qone(), x), qsub(qone(), y))); }

// correctSumAsPolyInFirstVariable (gkr_circuit.zig:390-429).
Poly correct_sum(qm31 f0, qm31 f2, qm31 claim, const std::vector<qm31>& y, uint32_t k) {
    SR_CHECK(k >= 1 && k <= y.size(), "gkr: invalid k");
    const size_t n = y.size();
    qm31 eq_prefix = qone();
    for (size_t i = 0; i < n - k + 1; i++) eq_prefix = qmul(eq_prefix, qsub(qone(), y[i]));
    SR_CHECK(!qiszero(eq_prefix), "gkr: division by zero");
    const qm31 a_const = qinv(eq_prefix);
    const qm31 yk = y[n - k];
    const qm31 den = qsub(qone(), qadd(yk, yk));
    SR_CHECK(!qiszero(den), "gkr: division by zero");
    const qm31 b_const = qdiv(qsub(qone(), yk), den);
    const qm31 two = qfromm(2);
    const qm31 r0 = qmul(qmul(f0, qsub(qone(), yk)), a_const);
    const qm31 r2 = qmul(qmul(f2, eq_one(two, yk)), a_const);
    const qm31 xs[4] = {qzero(), qone(), two, b_const};
    const qm31 ys[4] = {r0, qsub(claim, r0), r2, qzero()};
    return lagrange(xs, ys);
}

qm31 from_acc(const uint64_t* a) { return {{mred64(a[0]), mred64(a[1])}, {mred64(a[2]), mred64(a[3])}}; }

}  // namespace

// GkrMultivariatePolyOracle.sumAsPolyInFirstVariable (gkr_circuit.zig:261-316) for one
// layer on the device: returns the trimmed coefficients of the round polynomial.
std::vector<qm31> gkr_sum_as_poly(Ctx& c, int kind, const uint32_t* d_eq, const uint32_t* d_num,
                                  const uint32_t* d_den, uint32_t n_vars, qm31 lambda, qm31 claim,
                                  const std::vector<qm31>& y, qm31 correction) {
    SR_CHECK(n_vars >= 2, "gkr: the oracle needs at least one variable");
    SR_CHECK(n_vars - 1 <= y.size(), "gkr: eq point shorter than the oracle");
    DBuf acc(&c, 64);
    launch_zero_u64(c, acc.as<unsigned long long>(), 8);
    launch_gkr_sums(c, kind, d_eq, d_num, d_den, n_vars, lambda, acc.as<unsigned long long>());
    uint64_t* h = (uint64_t*)c.pinned(64);
    SR_HIP(hipMemcpyAsync(h, acc.p, 64, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    const qm31 f0 = qmul(from_acc(h), correction), f2 = qmul(from_acc(h + 4), correction);
    return correct_sum(f0, f2, claim, y, n_vars - 1).c;
}


// ====================================================================== device-resident batch prover
namespace {

constexpr uint32_t GD_MAX_INST = 8;
constexpr uint32_t GD_MAX_VARS = 30;
constexpr uint32_t GD_SMALL_LOG = 9;  // rounds on layers of <= 2^9 values run in one workgroup (measured best with 10 within 2%)
constexpr uint32_t GD_FT = 512;       // one-workgroup kernel: 7 summing waves + 1 for the pending update
constexpr uint32_t GD_RT = 256;       // threads per workgroup of the round and begin kernels
constexpr uint32_t GD_REC_ROUND = 17;  // [len, 4 QM31] per sum-check round
constexpr uint32_t GD_REC_MASK = 16;   // [num0, num1, den0, den1] per instance and layer

// Layer stack of one instance: layer l has l variables (2^l values); num is QM31 AoS,
// M31 for a multiplicities input layer, absent for grand products and singles.
struct GdInst {
    const uint32_t* num[GD_MAX_VARS + 1];
    const uint32_t* den[GD_MAX_VARS + 1];
    uint32_t kind[GD_MAX_VARS + 1];
    uint32_t n_vars;
    uint32_t* scr_num;  // fixed layers: 2^(n_vars-1) QM31 each
    uint32_t* scr_den;
};

struct GdState {
    uint32_t dg[8];
    uint32_t n_draws, status, pad0, pad00;
    uint32_t outputs[GD_MAX_INST][2][4];  // transcript order: [den] or [num, den]
    uint32_t claims[GD_MAX_INST][2][4];
    uint32_t sc_claim[GD_MAX_INST][4];
    uint32_t corr[GD_MAX_INST][4];
    uint32_t alpha_pow[GD_MAX_INST][4];
    uint32_t lambda[4];
    uint32_t r[4];
    uint32_t ood[GD_MAX_VARS + 1][4];  // eq point of the current layer
    uint32_t asg[GD_MAX_VARS][4];      // challenges of the current layer's rounds
    uint32_t pre[GD_MAX_VARS][11][4];  // per round: F0, F2, basis rows 1..3 (3 each)
    uint32_t mask[GD_MAX_INST][4][4];
    // the last round's claim update, applied while the next round sums (or dropped at the
    // layer end, where the claims are rebuilt): its coefficients, round and instance flags
    uint32_t craw[GD_MAX_INST][4][4];
    uint32_t pend, pend_rho, pad1, pad2;
    uint32_t pend_flags[GD_MAX_INST];  // bit 0: active, bit 1: summed in that round
    // the current round's coefficient c_q of instance i as A * s0 + C * s2 + B in the round
    // sums s0, s2 (set at the round start from the claim, the correction and the round's
    // basis rows), and the instance's flags in this round (bit 0: active, bit 1: summed)
    uint32_t fin[GD_MAX_INST][4][3][4];
    uint32_t rflags[GD_MAX_INST];
};
static_assert(sizeof(GdState) % 8 == 0, "state copy in 8-byte words");
constexpr uint32_t GD_MAX_BX = 1024;   // workgroups per instance in a round launch
constexpr uint32_t GD_CNT_GROUPS = 16;  // completion sub-counters, one 128-byte line each
constexpr uint32_t GD_CNT_STRIDE = 32;  // words between counters

__device__ __forceinline__ qm31 gd_ld(const uint32_t* p, uint64_t i) {
    const uint4 v = ((const uint4*)p)[i];
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void gd_st(uint32_t* p, uint64_t i, qm31 v) {
    ((uint4*)p)[i] = make_uint4(v.c0.a, v.c0.b, v.c1.a, v.c1.b);
}
__device__ __forceinline__ qm31 gd_num(uint32_t kind, const uint32_t* num, uint64_t j) {
    if (kind == GKR_LOGUP_GENERIC) return gd_ld(num, j);
    if (kind == GKR_LOGUP_MULTIPLICITIES) return qfromm(num[j]);
    return qone();
}
__device__ __forceinline__ uint32_t gd_fixed_kind(uint32_t kind) {
    return kind == GKR_LOGUP_MULTIPLICITIES ? (uint32_t)GKR_LOGUP_GENERIC : kind;
}
__device__ __forceinline__ bool gd_has_num(uint32_t kind) {
    return kind == GKR_LOGUP_GENERIC || kind == GKR_LOGUP_MULTIPLICITIES;
}
__device__ __forceinline__ qm31 gd_shfl(qm31 v, int src) {
    return {{(uint32_t)__shfl((int)v.c0.a, src), (uint32_t)__shfl((int)v.c0.b, src)},
            {(uint32_t)__shfl((int)v.c1.a, src), (uint32_t)__shfl((int)v.c1.b, src)}};
}
__device__ __forceinline__ qm31 gd_shfl_xor(qm31 v, int m) {
    return {{(uint32_t)__shfl_xor((int)v.c0.a, m), (uint32_t)__shfl_xor((int)v.c0.b, m)},
            {(uint32_t)__shfl_xor((int)v.c1.a, m), (uint32_t)__shfl_xor((int)v.c1.b, m)}};
}
template <int CTRL>
__device__ __forceinline__ qm31 gd_qperm(qm31 v) {
    return {{quad_perm<CTRL>(v.c0.a), quad_perm<CTRL>(v.c0.b)}, {quad_perm<CTRL>(v.c1.a), quad_perm<CTRL>(v.c1.b)}};
}
// The 8 sums of a wave by recursive halving: at each exchange a lane keeps half of its values
// and adds its partner's copy of that half (lane bit 5 picks the half first, then bits 4 and
// 3), then three full exchanges.  6 dependent rounds of word exchanges where 8 separate
// reductions take 48; the rounds sit on the latency path of every round.  Afterwards lane l
// holds the total of value (l >> 3) & 7.
__device__ __forceinline__ uint64_t gd_xchg(uint64_t v, uint32_t addr) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute((int)addr, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t gd_wave_sum8(uint64_t (&x)[8], uint32_t lane) {
#pragma unroll
    for (int st = 0; st < 3; st++) {
        const uint32_t h = 4u >> st, b = 32u >> st;
        // the halves are picked with a lane mask, not a select of two elements (which the
        // compiler turns into an indexed read of the whole array)
        const uint64_t m = 0ull - (uint64_t)((lane & b) != 0);  // all ones: keep the upper half
        const uint32_t addr = (lane ^ b) << 2;
#pragma unroll
        for (uint32_t i = 0; i < h; i++) {
            const uint64_t lo = x[i], hi = x[i + h];
            const uint64_t send = (lo & m) | (hi & ~m);
            x[i] = (lo ^ hi ^ send) + gd_xchg(send, addr);
        }
    }
    uint64_t v = x[0];
    v += gd_xchg(v, (lane ^ 4) << 2);
    v += gd_xchg(v, (lane ^ 2) << 2);
    v += gd_xchg(v, (lane ^ 1) << 2);
    return v;
}
__device__ __forceinline__ void gd_qstore(uint32_t* w, qm31 v) { qstore(v, w); }

// Instance view at GKR layer L: active once n_vars + L >= n_layers; its oracle layer has
// onv + 1 variables, onv = L + n_vars - n_layers.
struct GdView {
    bool active;
    uint32_t onv;
};
__device__ __forceinline__ GdView gd_view(const GdInst& in, uint32_t n_layers, uint32_t L) {
    GdView v;
    v.active = in.n_vars + L >= n_layers;
    v.onv = v.active ? in.n_vars + L - n_layers : 0;
    return v;
}

// Sum-check term k of one instance at round rho (nt = 2^(rem-1) terms): the four values
// of the current layer at rows 2k, 2k+1, 2(nt+k), 2(nt+k)+1, fixed on the fly from the
// previous layer (2x as long) when `fix`, then the terms at x = 0 and x = 2
// (gkr_circuit.zig:505-643), each multiplied by eq[k].
// The term is spread over a quad (all four lanes active, same k): lane j fixes and
// stores value j, the four values are exchanged with DPP, and every lane forms one
// product pair f * (A*B + C*D) chosen by its lane index (branch-free); lanes 0 and 2 end
// with t0 and t2.  Returns e * t in lanes 0 (x = 0) and 2 (x = 2), zero elsewhere.
__device__ __forceinline__ qm31 gd_term_quad(bool fix, uint32_t kind, const uint32_t* num, const uint32_t* den,
                                             uint32_t* dnum, uint32_t* dden, qm31 r, uint64_t nt, uint64_t k,
                                             qm31 lambda, const uint32_t* eq, uint32_t j) {
    const uint64_t jj = j < 2 ? 2 * k + j : 2 * (nt + k) + (j - 2);
    const uint64_t h = 4 * nt;
    const bool has_num = gd_has_num(kind);
    const qm31 e = gd_ld(eq, k);
    const qm31 da = gd_ld(den, jj);
    const qm31 db = fix ? gd_ld(den, jj + h) : qzero();
    const qm31 na = has_num ? gd_num(kind, num, jj) : qone();
    const qm31 nb = has_num && fix ? gd_num(kind, num, jj + h) : qone();
    const qm31 d = fix ? qadd(qmul(qsub(db, da), r), da) : da;
    const qm31 n = has_num && fix ? qadd(qmul(qsub(nb, na), r), na) : na;
    if (fix) {
        gd_st(dden, jj, d);
        if (has_num) gd_st(dnum, jj, n);
    }
    const qm31 d0 = quad_bcast_q<0>(d), d1 = quad_bcast_q<1>(d), d2 = quad_bcast_q<2>(d), d3 = quad_bcast_q<3>(d);
    const qm31 d20 = qsub(qadd(d2, d2), d0), d21 = qsub(qadd(d3, d3), d1);
    const bool l0 = j == 0, l1 = j == 1, l2 = j == 2;
    qm31 x;
    if (kind == GKR_GRAND_PRODUCT) {
        // lane 0: d0 d1 = t0; lane 2: d20 d21 = t2
        const qm31 A = l0 ? d0 : (l2 ? d20 : qzero());
        const qm31 B = l0 ? d1 : d21;
        x = qmul(A, B);
    } else {
        const qm31 n0 = quad_bcast_q<0>(n), n1 = quad_bcast_q<1>(n), n2 = quad_bcast_q<2>(n), n3 = quad_bcast_q<3>(n);
        const qm31 n20 = qsub(qadd(n2, n2), n0), n21 = qsub(qadd(n3, n3), n1);
        // t0 = (d1 n0 + d0 n1) + lambda d0 d1,  t2 = (d21 n20 + d20 n21) + lambda d20 d21
        const qm31 A = l0 ? d1 : (l1 ? d0 : (l2 ? d21 : d20));
        const qm31 B = l0 ? n0 : (l1 ? d1 : (l2 ? n20 : d21));
        const qm31 C = l0 ? d0 : (l2 ? d20 : qzero());
        const qm31 D = l0 ? n1 : n21;
        x = qadd(qmul(A, B), qmul(C, D));
        x = qmul((j & 1) ? lambda : qone(), x);
        x = qadd(x, gd_qperm<QP_ROT1>(x));
    }
    return (l0 || l2) ? qmul(e, x) : qzero();
}

// Where instance `in` reads round rho of layer L from: the stack layer (no fix yet this
// layer), or the fix scratch; `fix` = the previous round's variable is fixed on the fly.
struct GdSrc {
    bool fix;
    uint32_t kind;  // kind of the values this round sums over
    uint32_t src_kind;
    const uint32_t* num;
    const uint32_t* den;
};
__device__ __forceinline__ GdSrc gd_src(const GdInst& in, uint32_t onv, uint32_t L, uint32_t rho) {
    const uint32_t lv = onv + 1;
    const uint32_t fixes = rho - (L - onv);  // fixes done before this round's sums
    GdSrc s;
    s.fix = fixes >= 1;
    const bool from_stack = fixes <= 1;
    s.src_kind = from_stack ? in.kind[lv] : gd_fixed_kind(in.kind[lv]);
    s.kind = s.fix ? gd_fixed_kind(s.src_kind) : s.src_kind;
    s.num = from_stack ? in.num[lv] : in.scr_num;
    s.den = from_stack ? in.den[lv] : in.scr_den;
    return s;
}

// Round start: the previous round's pending claim update (sc_claim_i = rp_i(r) and, for the
// instances summed in that round, corr_i *= eq(r, y_rho)), then the finishing coefficients
// of this round.  With y0 = s0 corr F0, y2 = s2 corr F2 and
// y1 = claim - y0, the round polynomial's coefficient q of an instance summed in the round is
// y0 for q = 0 and b1 y0 + b2 y1 + b3 y2 = (b1 - b2) y0 + b2 claim + b3 y2 for q = 1..3 (basis
// rows b of the round), so A, B, C depend only on values known before the sums.  One full
// wave; lane t handles instance t>>2, coefficient t&3; the quad's lanes share the instance's
// claim and correction in registers.
__device__ __forceinline__ void gd_round_prep(GdState* S, const GdInst* I, uint32_t n_inst, uint32_t n_layers,
                                              uint32_t L, uint32_t rho) {
    const uint32_t t = threadIdx.x & 63;
    const uint32_t i = t >> 2, q = t & 3;
    const bool in_range = i < n_inst;
    qm31 sc = in_range ? qload(S->sc_claim[i]) : qzero();
    qm31 corr = in_range ? qload(S->corr[i]) : qzero();
    if (S->pend) {
        const qm31 r = qload(S->r);
        const qm31 craw = in_range ? qload(S->craw[i][q]) : qzero();
        qm31 pw = qone();
        if (q >= 1) pw = r;
        if (q >= 2) pw = qmul(pw, r);
        if (q >= 3) pw = qmul(pw, r);
        qm31 v = qmul(craw, pw);
        v = qadd(v, gd_qperm<QP_ROT2>(v));
        v = qadd(v, gd_qperm<QP_ROT1>(v));  // every lane of the quad: rp_i(r)
        const uint32_t fl = in_range ? S->pend_flags[i] : 0u;
        if (fl & 1) {
            sc = v;
            if (fl & 2) {
                const qm31 y = qload(S->ood[S->pend_rho]);
                const qm31 e = qadd(qmul(r, y), qmul(qsub(qone(), r), qsub(qone(), y)));
                corr = qmul(corr, e);
            }
            if (q == 0) {
                gd_qstore(S->sc_claim[i], sc);
                if (fl & 2) gd_qstore(S->corr[i], corr);
            }
        }
    }
    if (!in_range) return;
    const GdView vw = gd_view(I[i], n_layers, L);
    const bool act = vw.active, part = act && vw.onv >= L - rho;
    if (q == 0) S->rflags[i] = (act ? 1u : 0u) | (part ? 2u : 0u);
    if (!part) return;
    const uint32_t(*pr)[4] = S->pre[rho];
    const qm31 c0 = qmul(corr, qload(pr[0])), c2 = qmul(corr, qload(pr[1]));
    qm31 A, B, C;
    if (q == 0) {
        A = c0;
        B = qzero();
        C = qzero();
    } else {
        const uint32_t b = 2 + 3 * (q - 1);
        const qm31 b1 = qload(pr[b]), b2 = qload(pr[b + 1]), b3 = qload(pr[b + 2]);
        A = qmul(qsub(b1, b2), c0);
        B = qmul(b2, sc);
        C = qmul(b3, c2);
    }
    gd_qstore(S->fin[i][q][0], A);
    gd_qstore(S->fin[i][q][1], B);
    gd_qstore(S->fin[i][q][2], C);
}

// Turns the round sums (sums[8 i + c]: r(0) then r(2) coordinates before the eq
// correction, 64-bit) into the batched round polynomial, mixes it and draws the round
// challenge; the claim update is left pending.  Wave 0 only, with the previous round's
// update already applied; lane t handles instance t>>2, coefficient t&3.
__device__ __forceinline__ void gd_finish(GdState* S, const GdInst* I, uint32_t n_inst, uint32_t n_layers, uint32_t L, uint32_t rho,
                          const unsigned long long* sums, uint32_t* rec) {
    const uint32_t t = threadIdx.x;
    const uint32_t i = t >> 2, q = t & 3;
    const uint32_t rem = L - rho;
    (void)I;
    (void)n_layers;
    (void)rem;
    qm31 craw = qzero();
    if (i < n_inst) {
        const uint32_t fl = S->rflags[i];  // set by gd_round_prep for this round
        const bool act = fl & 1, part = fl & 2;
        if (part) {
            const unsigned long long* a = sums + 8 * i;
            const qm31 s0 = {{mred64(a[0]), mred64(a[1])}, {mred64(a[2]), mred64(a[3])}};
            const qm31 s2 = {{mred64(a[4]), mred64(a[5])}, {mred64(a[6]), mred64(a[7])}};
            craw = qadd(qadd(qmul(qload(S->fin[i][q][0]), s0), qmul(qload(S->fin[i][q][2]), s2)),
                        qload(S->fin[i][q][1]));
        } else if (act && q == 0) {
            craw = qmulm(qload(S->sc_claim[i]), (P + 1) / 2);
        }
        gd_qstore(S->craw[i][q], craw);
        if (q == 0) S->pend_flags[i] = fl;
    }
    qm31 cv = i < n_inst ? qmul(craw, qload(S->alpha_pow[i])) : qzero();
#pragma unroll
    for (int m = 4; m < 32; m <<= 1) cv = qadd(cv, gd_shfl_xor(cv, m));
    const uint64_t nzmask = __ballot(t < 4 && !qiszero(cv)) & 0xFull;
    const uint32_t len = nzmask ? 64 - __clzll(nzmask) : 0;
    if (t < 4) {
        gd_qstore(rec + 1 + 4 * t, cv);
        if (t == 0) rec[0] = len;
        const QuadLane Lq(t);
        uint32_t dg[8];
#pragma unroll
        for (int k = 0; k < 8; k++) dg[k] = S->dg[k];
        uint32_t pay[16];
        qstore(quad_bcast_q<0>(cv), pay);
        qstore(quad_bcast_q<1>(cv), pay + 4);
        qstore(quad_bcast_q<2>(cv), pay + 8);
        qstore(quad_bcast_q<3>(cv), pay + 12);
        qchan_mix(Lq, dg, pay, 4 * len);
        uint32_t nd = 0, st = 0;
        const qm31 r = qchan_draw(Lq, dg, nd, st);
        if (t == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) S->dg[k] = dg[k];
            S->n_draws = nd;
            if (st) S->status = 1;
            gd_qstore(S->r, r);
            gd_qstore(S->asg[rho], r);
            S->pend = 1;
            S->pend_rho = rho;
        }
    }
}

// Layer begin (all threads of the block call it; it synchronises the block).
__device__ __forceinline__ void gd_begin(GdState* S, const GdInst* I, uint32_t n_inst, uint32_t n_layers, uint32_t L) {
    const uint32_t t = threadIdx.x;
    const uint32_t remaining = n_layers - L;
    if (t < 4) {
        const QuadLane Lq(t);
        uint32_t dg[8];
#pragma unroll
        for (int k = 0; k < 8; k++) dg[k] = S->dg[k];
        uint32_t nd = S->n_draws, st = 0;
        for (uint32_t i = 0; i < n_inst; i++) {
            if (!gd_view(I[i], n_layers, L).active) continue;
            const uint32_t nc = I[i].kind[0] == GKR_GRAND_PRODUCT ? 1 : 2;
            const bool fresh = I[i].n_vars == remaining;
            uint32_t pay[16];
#pragma unroll
            for (int k = 0; k < 8; k++) pay[k] = fresh ? (&S->outputs[i][0][0])[k] : (&S->claims[i][0][0])[k];
#pragma unroll
            for (int k = 8; k < 16; k++) pay[k] = 0;
            if (fresh && t == 0)
                for (int k = 0; k < 8; k++) (&S->claims[i][0][0])[k] = pay[k];
            qchan_mix(Lq, dg, pay, 4 * nc);
            nd = 0;
        }
        const qm31 alpha = qchan_draw(Lq, dg, nd, st);
        const qm31 lambda = qchan_draw(Lq, dg, nd, st);
        if (t == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) S->dg[k] = dg[k];
            S->n_draws = nd;
            if (st) S->status = 1;
            S->pend = 0;
            gd_qstore(S->lambda, lambda);
            // the batch combines the active instances only, in order (sumcheck.zig:215-238)
            qm31 ap = qone();
            for (uint32_t i = 0; i < n_inst; i++) {
                const bool a = gd_view(I[i], n_layers, L).active;
                gd_qstore(S->alpha_pow[i], a ? ap : qzero());
                if (a) ap = qmul(ap, alpha);
            }
        }
    } else if (t >= 64 && t < 64 + L) {
        // round rho's constants from y = ood[0..L)
        const uint32_t rho = t - 64;
        qm31 pp = qone();
        for (uint32_t k = 0; k <= rho; k++) pp = qmul(pp, qsub(qone(), qload(S->ood[k])));
        const qm31 y = qload(S->ood[rho]);
        const qm31 dd = qsub(qone(), qadd(y, y));
        const qm31 pd = qmul(pp, dd);
        const qm31 ipd = qinv(pd);
        const qm31 ipp = qmul(ipd, dd), idd = qmul(ipd, pp);
        const qm31 one = qone(), two = qfromm(2), three = qfromm(3);
        const qm31 b = qmul(qsub(one, y), idd);
        const qm31 bm1 = qsub(b, one), tmb = qsub(two, b);
        const qm31 qq = qmul(qmul(b, bm1), tmb);
        const qm31 iq = qinv(qq);
        const qm31 ib = qmul(qmul(bm1, tmb), iq), ibm1 = qmul(qmul(b, tmb), iq), itmb = qmul(qmul(b, bm1), iq);
        const uint32_t inv2 = (P + 1) / 2;
        const qm31 s0 = qneg(qmulm(ib, inv2)), s1 = ibm1, s2 = qmulm(itmb, inv2);
        uint32_t(*pr)[4] = S->pre[rho];
        gd_qstore(pr[0], qmul(qsub(one, y), ipp));                          // F0 = (1-y) / prefix
        gd_qstore(pr[1], qmul(qsub(qmul(three, y), one), ipp));            // F2 = eq(2, y) / prefix
        gd_qstore(pr[2], qmul(s0, qadd(two, qmul(three, b))));            // x^1
        gd_qstore(pr[3], qmul(s1, qadd(b, b)));
        gd_qstore(pr[4], qmul(s2, b));
        gd_qstore(pr[5], qneg(qmul(s0, qadd(three, b))));                 // x^2
        gd_qstore(pr[6], qneg(qmul(s1, qadd(two, b))));
        gd_qstore(pr[7], qneg(qmul(s2, qadd(one, b))));
        gd_qstore(pr[8], s0);                                               // x^3
        gd_qstore(pr[9], s1);
        gd_qstore(pr[10], s2);
        if (qiszero(pd) || qiszero(qq)) S->status = 2;
    }
    __syncthreads();
    if (t < n_inst) {
        const GdView v = gd_view(I[t], n_layers, L);
        if (v.active) {
            const uint32_t nc = I[t].kind[0] == GKR_GRAND_PRODUCT ? 1 : 2;
            const qm31 lambda = qload(S->lambda);
            qm31 cl = qzero();
            for (uint32_t c = nc; c > 0; c--) cl = qadd(qmul(cl, lambda), qload(S->claims[t][c - 1]));
            gd_qstore(S->sc_claim[t], qmulm(cl, 1u << (L - v.onv)));
            gd_qstore(S->corr[t], qone());
        }
    }
    __syncthreads();
}

// eq table of layer L: eq[i] = (1 - y_0) prod_{j < L-1} (bit_{L-2-j}(i) ? y_{j+1} : 1 - y_{j+1})
// (EqEvals.generate(ood) = genEqEvals(ood[1..], 1 - ood[0]), gkr_circuit.zig:651-674).
__device__ __forceinline__ void gd_eq_entry(const GdState* S, uint32_t L, uint64_t i, uint32_t* eq) {
    qm31 acc = qsub(qone(), qload(S->ood[0]));
    for (uint32_t j = 0; j + 1 < L; j++) {
        const qm31 y = qload(S->ood[j + 1]);
        acc = qmul(acc, ((i >> (L - 2 - j)) & 1) ? y : qsub(qone(), y));
    }
    gd_st(eq, i, acc);
}

// Layer end: masks of every active instance, mixed in order; challenge; next claims and
// eq point (gkr_prover.zig:232-276, gkr_circuit.zig:341-388).
__device__ __forceinline__ void gd_end(GdState* S, const GdInst* I, uint32_t n_inst, uint32_t n_layers, uint32_t L, uint32_t* rec_mask) {
    const uint32_t t = threadIdx.x;
    if (t < n_inst) {
        const GdView v = gd_view(I[t], n_layers, L);
        if (v.active) {
            const GdInst& in = I[t];
            const uint32_t lv = v.onv + 1;
            qm31 n0, n1, d0, d1;
            if (v.onv == 0) {
                const uint32_t k = in.kind[lv];
                d0 = gd_ld(in.den[lv], 0);
                d1 = gd_ld(in.den[lv], 1);
                n0 = k == GKR_GRAND_PRODUCT ? qzero() : gd_num(k, in.num[lv], 0);
                n1 = k == GKR_GRAND_PRODUCT ? qzero() : gd_num(k, in.num[lv], 1);
            } else {
                const qm31 r = qload(S->r);
                const bool stack = v.onv == 1;
                const uint32_t k = stack ? in.kind[lv] : gd_fixed_kind(in.kind[lv]);
                const uint32_t* dn = stack ? in.den[lv] : in.scr_den;
                const uint32_t* nn = stack ? in.num[lv] : in.scr_num;
                qm31 a = gd_ld(dn, 0), b = gd_ld(dn, 2);
                d0 = qadd(qmul(r, qsub(b, a)), a);
                a = gd_ld(dn, 1), b = gd_ld(dn, 3);
                d1 = qadd(qmul(r, qsub(b, a)), a);
                if (gd_has_num(k)) {
                    a = gd_num(k, nn, 0), b = gd_num(k, nn, 2);
                    n0 = qadd(qmul(r, qsub(b, a)), a);
                    a = gd_num(k, nn, 1), b = gd_num(k, nn, 3);
                    n1 = qadd(qmul(r, qsub(b, a)), a);
                } else {
                    n0 = k == GKR_GRAND_PRODUCT ? qzero() : qone();
                    n1 = n0;
                }
            }
            gd_qstore(S->mask[t][0], n0);
            gd_qstore(S->mask[t][1], n1);
            gd_qstore(S->mask[t][2], d0);
            gd_qstore(S->mask[t][3], d1);
            uint32_t* rm = rec_mask + (uint64_t)t * GD_REC_MASK;
            gd_qstore(rm, n0);
            gd_qstore(rm + 4, n1);
            gd_qstore(rm + 8, d0);
            gd_qstore(rm + 12, d1);
        }
    }
    __syncthreads();
    if (t < 4) {
        const QuadLane Lq(t);
        uint32_t dg[8];
#pragma unroll
        for (int k = 0; k < 8; k++) dg[k] = S->dg[k];
        uint32_t nd = S->n_draws, st = 0;
        for (uint32_t i = 0; i < n_inst; i++) {
            if (!gd_view(I[i], n_layers, L).active) continue;
            const bool gp = I[i].kind[0] == GKR_GRAND_PRODUCT;
            uint32_t pay[16];
#pragma unroll
            for (int k = 0; k < 16; k++) pay[k] = gp ? (k < 8 ? (&S->mask[i][2][0])[k] : 0u) : (&S->mask[i][0][0])[k];
            qchan_mix(Lq, dg, pay, gp ? 8 : 16);
            nd = 0;
        }
        const qm31 ch = qchan_draw(Lq, dg, nd, st);
        if (t == 0) {
#pragma unroll
            for (int k = 0; k < 8; k++) S->dg[k] = dg[k];
            S->n_draws = nd;
            if (st) S->status = 1;
            S->pend = 0;  // the last round's claim update is not needed: the claims come from the masks
            for (uint32_t i = 0; i < n_inst; i++) {
                if (!gd_view(I[i], n_layers, L).active) continue;
                const bool gp = I[i].kind[0] == GKR_GRAND_PRODUCT;
                for (uint32_t c = 0; c < (gp ? 1u : 2u); c++) {
                    const uint32_t col = gp ? 1 : c;  // mask column: den (GP) or num, den
                    const qm31 m0 = qload(S->mask[i][2 * col]), m1 = qload(S->mask[i][2 * col + 1]);
                    gd_qstore(S->claims[i][c], qadd(qmul(ch, qsub(m1, m0)), m0));
                }
            }
            for (uint32_t j = 0; j < L; j++)
                for (int k = 0; k < 4; k++) S->ood[j][k] = S->asg[j][k];
            gd_qstore(S->ood[L], ch);
        }
    }
    __syncthreads();
}

// One workgroup: whole layers [L0, L1), the first one starting at round rho0 (its begin
// already done when rho0 > 0); every layer ends here.  The state lives in LDS meanwhile.
__global__ __launch_bounds__(GD_FT) void gkr_fused_kernel(GdState* __restrict__ gS, const GdInst* __restrict__ I,
                                                          uint32_t n_inst, uint32_t n_layers, uint32_t L0, uint32_t L1,
                                                          uint32_t rho0, uint32_t* __restrict__ eq,
                                                          uint32_t* __restrict__ rec) {
    __shared__ GdState S;
    // round sums, double-buffered by round parity: the pending-update wave clears the other
    // buffer while a round accumulates into its own (LDS atomics, one add per term and word)
    __shared__ unsigned long long red[2][GD_MAX_INST][8];
    const uint32_t t = threadIdx.x;
    const uint32_t wv = t >> 6, ln = t & 63;
    {
        const uint64_t* src = (const uint64_t*)gS;
        uint64_t* dst = (uint64_t*)&S;
        for (uint32_t k = t; k < sizeof(GdState) / 8; k += GD_FT) dst[k] = src[k];
        for (uint32_t k = t; k < 2 * GD_MAX_INST * 8; k += GD_FT) (&red[0][0][0])[k] = 0;
    }
    __syncthreads();
    if (L0 == 0 && rho0 == 0 && t < n_inst) {
        // output layer (gkr_circuit.zig:162-190)
        const GdInst& in = I[t];
        const qm31 d = gd_ld(in.den[0], 0);
        if (in.kind[0] == GKR_GRAND_PRODUCT) {
            gd_qstore(S.outputs[t][0], d);
        } else {
            gd_qstore(S.outputs[t][0], gd_num(in.kind[0], in.num[0], 0));
            gd_qstore(S.outputs[t][1], d);
        }
    }
    __syncthreads();
    const uint32_t total_rounds = n_layers * (n_layers - 1) / 2;
    uint32_t rc = 0;  // rounds run by this launch (sum buffer parity)
    for (uint32_t L = L0; L < L1; L++) {
        const uint32_t rb = L == L0 ? rho0 : 0;
        if (rb == 0) {
            gd_begin(&S, I, n_inst, n_layers, L);
            if (L >= 1)
                for (uint32_t k = t; k < (1u << (L - 1)); k += GD_FT) gd_eq_entry(&S, L, k, eq);
            __syncthreads();
        }
        for (uint32_t rho = rb; rho < L; rho++, rc++) {
            const uint32_t rem = L - rho;
            const uint64_t nt = 1ull << (rem - 1);
            const uint32_t b = rc & 1;
            if (wv == 7) {
                gd_round_prep(&S, I, n_inst, n_layers, L, rho);
                (&red[b ^ 1][0][0])[ln] = 0;
            } else {
                // (instance, term) pairs over the 112 quads of waves 0..6; pair x = p * nt + k for
                // the p-th instance summed in this round
                const qm31 lambda = qload(S.lambda), r = qload(S.r);
                uint32_t n_part = 0;
                for (uint32_t i = 0; i < n_inst; i++) {
                    const GdView v = gd_view(I[i], n_layers, L);
                    n_part += v.active && v.onv >= rem;
                }
                const uint32_t j = t & 3;
                for (uint64_t x = t >> 2; x < (uint64_t)n_part * nt; x += 7 * 16) {
                    const uint32_t p = (uint32_t)(x >> (rem - 1));
                    const uint64_t k = x & (nt - 1);
                    uint32_t i = 0;
                    for (uint32_t ii = 0, cnt = 0; ii < n_inst; ii++) {
                        const GdView v = gd_view(I[ii], n_layers, L);
                        if (!(v.active && v.onv >= rem)) continue;
                        if (cnt++ == p) {
                            i = ii;
                            break;
                        }
                    }
                    const GdView v = gd_view(I[i], n_layers, L);
                    const GdSrc s = gd_src(I[i], v.onv, L, rho);
                    const qm31 y = gd_term_quad(s.fix, s.src_kind, s.num, s.den, I[i].scr_num, I[i].scr_den, r, nt, k,
                                                lambda, eq, j);
                    if (j == 0 || j == 2) {
                        unsigned long long* dst = &red[b][i][j == 0 ? 0 : 4];
                        atomicAdd(dst + 0, (unsigned long long)y.c0.a);
                        atomicAdd(dst + 1, (unsigned long long)y.c0.b);
                        atomicAdd(dst + 2, (unsigned long long)y.c1.a);
                        atomicAdd(dst + 3, (unsigned long long)y.c1.b);
                    }
                }
            }
            __syncthreads();
            if (t < 64)
                gd_finish(&S, I, n_inst, n_layers, L, rho, &red[b][0][0],
                          rec + (uint64_t)(L * (L - 1) / 2 + rho) * GD_REC_ROUND);
            __syncthreads();
        }
        gd_end(&S, I, n_inst, n_layers, L,
               rec + (uint64_t)total_rounds * GD_REC_ROUND + (uint64_t)L * n_inst * GD_REC_MASK);
    }
    {
        const uint64_t* src = (const uint64_t*)&S;
        uint64_t* dst = (uint64_t*)gS;
        for (uint32_t k = t; k < sizeof(GdState) / 8; k += GD_FT) dst[k] = src[k];
    }
}

// Layer begin of a large layer plus its eq table over many workgroups (block 0 also runs the begin).
__global__ __launch_bounds__(GD_RT) void gkr_begin_eq_kernel(GdState* __restrict__ S, const GdInst* __restrict__ I,
                                                             uint32_t n_inst, uint32_t n_layers, uint32_t L,
                                                             uint32_t* __restrict__ eq) {
    const uint64_t k = (uint64_t)blockIdx.x * GD_RT + threadIdx.x;
    if (L >= 1 && k < (1ull << (L - 1))) gd_eq_entry(S, L, k, eq);
    if (blockIdx.x == 0) gd_begin(S, I, n_inst, n_layers, L);
}

// One large sum-check round: grid (x: terms, y: instance).  Every workgroup stores its
// sums in its own slot of `part` and counts itself done on one of 16 sub-counters (separate
// cache lines; the workgroup completing a sub-counter counts on the top counter), so no
// address takes more than ~nb/16 atomics; the last workgroup adds the slots up and finishes
// the round.  Workgroup (x = last, y = 0) sums nothing: it applies the previous round's pending
// claim update and sets this round's finishing coefficients (gd_round_prep) while the others
// sum; the release fence before its count (the only one) publishes them to the last workgroup.
__global__ __launch_bounds__(GD_RT) void gkr_round_kernel(GdState* __restrict__ S, const GdInst* __restrict__ I,
                                                          uint32_t n_inst, uint32_t n_layers, uint32_t L, uint32_t rho,
                                                          const uint32_t* __restrict__ eq, uint32_t* __restrict__ rec,
                                                          unsigned long long* __restrict__ part, uint32_t* __restrict__ cnt) {
    __shared__ unsigned long long red[GD_RT / 64][8];
    __shared__ unsigned long long sums[GD_MAX_INST][8];
    __shared__ uint32_t last;
    const uint32_t t = threadIdx.x;
    const uint32_t rem = L - rho;
    const uint64_t nt = 1ull << (rem - 1);
    // blockIdx.y = rank among the instances summed in this round
    uint32_t i = 0;
    for (uint32_t ii = 0, cnt_p = 0; ii < n_inst; ii++) {
        const GdView vi = gd_view(I[ii], n_layers, L);
        if (!(vi.active && vi.onv >= rem)) continue;
        if (cnt_p++ == blockIdx.y) {
            i = ii;
            break;
        }
    }
    // the last workgroup column runs the round-start prep (instance row 0) instead of sums, so
    // the prep overlaps the sums rather than delaying one summing workgroup
    const uint32_t nsb = gridDim.x - 1;
    const bool prep_blk = blockIdx.x == nsb;
    if (prep_blk && blockIdx.y == 0 && t < 64) gd_round_prep(S, I, n_inst, n_layers, L, rho);
    const GdView v = gd_view(I[i], n_layers, L);
    uint64_t a[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!prep_blk) {
        const GdSrc s = gd_src(I[i], v.onv, L, rho);
        const qm31 lambda = qload(S->lambda), r = qload(S->r);
        const uint32_t j = t & 3;
        for (uint64_t k = (uint64_t)blockIdx.x * (GD_RT / 4) + (t >> 2); k < nt; k += (uint64_t)nsb * (GD_RT / 4)) {
            const qm31 y = gd_term_quad(s.fix, s.src_kind, s.num, s.den, I[i].scr_num, I[i].scr_den, r, nt, k, lambda,
                                        eq, j);
            if (j == 0) {
                a[0] += y.c0.a, a[1] += y.c0.b, a[2] += y.c1.a, a[3] += y.c1.b;
            } else if (j == 2) {
                a[4] += y.c0.a, a[5] += y.c0.b, a[6] += y.c1.a, a[7] += y.c1.b;
            }
        }
    }
    {
        const uint64_t w = gd_wave_sum8(a, t & 63);
        if ((t & 7) == 0) red[t >> 6][(t & 63) >> 3] = w;
    }
    __syncthreads();
    // The slot is written with device-coherent stores (agent-scope atomic stores) by wave 0,
    // which waits for their acknowledgement before thread 0 counts the workgroup: no L2
    // writeback per workgroup.  The prep workgroup wrote the state with plain stores and
    // publishes it with one release; the last workgroup acquires.
    if (t < 8) {
        uint64_t w = 0;
        for (uint32_t k = 0; k < GD_RT / 64; k++) w += red[k][t];
        __hip_atomic_store(&part[((uint64_t)blockIdx.y * gridDim.x + blockIdx.x) * 8 + t], (unsigned long long)w,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (t < 64) {
        __builtin_amdgcn_s_waitcnt(0);
        if (t == 0) {
            if (prep_blk && blockIdx.y == 0) __threadfence();
            const uint32_t nb = gridDim.x * gridDim.y, b = blockIdx.y * gridDim.x + blockIdx.x;
            const uint32_t g = b % GD_CNT_GROUPS, ng = nb < GD_CNT_GROUPS ? nb : GD_CNT_GROUPS;
            const uint32_t gsize = (nb - g + GD_CNT_GROUPS - 1) / GD_CNT_GROUPS;
            uint32_t lst = 0;
            if (atomicAdd(&cnt[GD_CNT_STRIDE * g], 1u) == gsize - 1)
                lst = atomicAdd(&cnt[GD_CNT_STRIDE * GD_CNT_GROUPS], 1u) == ng - 1 ? 1u : 0u;
            last = lst;
        }
    }
    __syncthreads();
    if (!last) return;
    // acquire: the slots came with device-coherent stores and the prep's state with its
    // workgroup's release; this workgroup has nothing of its own to write back
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    if (t <= GD_CNT_GROUPS) atomicExch(&cnt[GD_CNT_STRIDE * t], 0u);
    for (uint32_t ii = 0, py = 0; ii < n_inst; ii++) {
        const GdView vi = gd_view(I[ii], n_layers, L);
        if (!(vi.active && vi.onv >= rem)) continue;
        uint64_t b8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t row = py++;
        for (uint32_t x = t; x < gridDim.x; x += GD_RT) {
            unsigned long long* pp = part + ((uint64_t)row * gridDim.x + x) * 8;
#pragma unroll
            for (int c = 0; c < 8; c++) b8[c] += __hip_atomic_load(&pp[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        {
            const uint64_t w = gd_wave_sum8(b8, t & 63);
            if ((t & 7) == 0) red[t >> 6][(t & 63) >> 3] = w;
        }
        __syncthreads();
        if (t < 8) {
            uint64_t w = 0;
            for (uint32_t k = 0; k < GD_RT / 64; k++) w += red[k][t];
            sums[ii][t] = w;
        }
        __syncthreads();
    }
    if (t < 64) gd_finish(S, I, n_inst, n_layers, L, rho, &sums[0][0], rec + (uint64_t)(L * (L - 1) / 2 + rho) * GD_REC_ROUND);
}

}  // namespace

// The state starts from a device channel (every other word zero) and hands its digest and
// draw counter back at the end; a failed draw or a zero division sets the channel status.
__global__ void gkr_chan_in_kernel(GdState* __restrict__ S, const DevChannel* __restrict__ ch) {
    uint32_t* w = (uint32_t*)S;
    constexpr uint32_t NW = sizeof(GdState) / 4;
    for (uint32_t i = threadIdx.x; i < NW; i += blockDim.x) {
        uint32_t v = 0;
        if (i < 8) v = ch->digest[i];
        else if (i == 8) v = ch->n_draws;
        w[i] = v;
    }
}
__global__ void gkr_chan_out_kernel(DevChannel* __restrict__ ch, const GdState* __restrict__ S) {
    const uint32_t t = threadIdx.x;
    if (t < 8) ch->digest[t] = S->dg[t];
    if (t == 8) ch->n_draws = S->n_draws;
    if (t == 9 && S->status != 0) ch->status = S->status;
}

struct GkrRun {
    std::vector<GkrInstance> inst;
    uint32_t n_inst = 0, n_layers = 0;
    uint64_t rec_words = 0;
    std::vector<DBuf> bufs;
    DBuf d_inst, d_state, d_rec, d_eq, d_part, d_cnt;
};

GkrRun* gkr_launch_on_channel(Ctx& c, DevChannel* d_ch, const std::vector<GkrInstance>& inst) {
    SR_CHECK(d_ch, "gkr: device channel");
    SR_CHECK(!inst.empty(), "gkr: empty batch");
    SR_CHECK(inst.size() <= GD_MAX_INST, "gkr: at most 8 instances per batch");
    std::unique_ptr<GkrRun> run(new GkrRun());
    run->inst = inst;
    const uint32_t n_inst = (uint32_t)inst.size();
    uint32_t n_layers = 0;
    for (auto& in : inst) {
        SR_CHECK(in.kind >= 0 && in.kind <= 3, "gkr: layer kind");
        SR_CHECK(in.n_vars >= 1 && in.n_vars <= GD_MAX_VARS, "gkr: an instance needs 1..30 variables");
        SR_CHECK(in.d_den && (in.kind == GKR_GRAND_PRODUCT || in.kind == GKR_LOGUP_SINGLES || in.d_num),
                 "gkr: missing layer columns");
        n_layers = std::max(n_layers, in.n_vars);
    }
    run->n_inst = n_inst;
    run->n_layers = n_layers;
    // Layer stacks (buildLayers, gkr_circuit.zig:431-447) and fix scratch.
    std::vector<GdInst> h_inst(n_inst);
    std::vector<DBuf>& bufs = run->bufs;
    for (uint32_t k = 0; k < n_inst; k++) {
        GdInst& g = h_inst[k];
        memset(&g, 0, sizeof(g));
        const uint32_t nv = inst[k].n_vars;
        const bool gp = inst[k].kind == GKR_GRAND_PRODUCT;
        g.n_vars = nv;
        for (uint32_t l = 0; l < nv; l++) g.kind[l] = gp ? GKR_GRAND_PRODUCT : GKR_LOGUP_GENERIC;
        g.kind[nv] = (uint32_t)inst[k].kind;
        g.num[nv] = inst[k].d_num;
        g.den[nv] = inst[k].d_den;
        for (uint32_t l = 0; l < nv; l++) {
            if (!gp) {
                bufs.emplace_back(&c, 16ull << l);
                g.num[l] = bufs.back().as<uint32_t>();
            }
            bufs.emplace_back(&c, 16ull << l);
            g.den[l] = bufs.back().as<uint32_t>();
        }
        for (int32_t l = (int32_t)nv - 1; l >= 0; l--)
            launch_gkr_next_layer(c, (int)g.kind[l + 1], g.num[l + 1], g.den[l + 1], (uint32_t)l + 1,
                                  (uint32_t*)g.num[l], (uint32_t*)g.den[l]);
        bufs.emplace_back(&c, 16ull << (nv - 1));
        g.scr_den = bufs.back().as<uint32_t>();
        // every LogUp stack layer below the input is generic (QM31 numerators), singles included
        if (!gp) {
            bufs.emplace_back(&c, 16ull << (nv - 1));
            g.scr_num = bufs.back().as<uint32_t>();
        }
    }
    run->d_inst = upload(c, h_inst.data(), h_inst.size());
    run->d_state = DBuf(&c, sizeof(GdState));
    hipLaunchKernelGGL(gkr_chan_in_kernel, dim3(1), dim3(256), 0, c.stream, run->d_state.as<GdState>(),
                       (const DevChannel*)d_ch);
    SR_HIP(hipGetLastError());
    const uint32_t total_rounds = n_layers * (n_layers - 1) / 2;
    run->rec_words = (uint64_t)total_rounds * GD_REC_ROUND + (uint64_t)n_layers * n_inst * GD_REC_MASK;
    run->d_rec = DBuf(&c, 4 * run->rec_words);
    run->d_eq = DBuf(&c, n_layers >= 2 ? (16ull << (n_layers - 2)) : 16);
    run->d_part = DBuf(&c, 64ull * n_inst * GD_MAX_BX);
    run->d_cnt = DBuf(&c, 4ull * GD_CNT_STRIDE * (GD_CNT_GROUPS + 1));
    launch_fill_u32(c, run->d_cnt.as<uint32_t>(), GD_CNT_STRIDE * (GD_CNT_GROUPS + 1), 0);
    GdState* S = run->d_state.as<GdState>();
    const GdInst* I = run->d_inst.as<GdInst>();
    uint32_t* eq = run->d_eq.as<uint32_t>();
    uint32_t* rec = run->d_rec.as<uint32_t>();
    // algorithmic bytes of one sum-check round: 96 B per term (four values read, two fixed
    // values written, one eq entry) for every instance summed in it
    auto round_bytes = [&](uint32_t L, uint32_t rho) {
        uint32_t n_part = 0;
        for (uint32_t k = 0; k < n_inst; k++)
            n_part += inst[k].n_vars + L >= n_layers && inst[k].n_vars + L - n_layers >= L - rho;
        return 96.0 * (double)(1ull << (L - rho - 1)) * n_part;
    };
    {
        double bytes = 0;
        for (uint32_t L = 1; L < std::min(n_layers, GD_SMALL_LOG); L++)
            for (uint32_t rho = 0; rho < L; rho++) bytes += round_bytes(L, rho);
        KTimer kt(c, "gkr_small_layers", bytes);
        hipLaunchKernelGGL(gkr_fused_kernel, dim3(1), dim3(GD_FT), 0, c.stream, S, I, n_inst, n_layers, 0u,
                           std::min(n_layers, GD_SMALL_LOG), 0u, eq, rec);
        SR_HIP(hipGetLastError());
    }
    for (uint32_t L = GD_SMALL_LOG; L < n_layers; L++) {
        {
            KTimer kt(c, "gkr_begin_eq", 16.0 * (double)(1ull << (L - 1)));
            const uint64_t blocks = ((1ull << (L - 1)) + GD_RT - 1) / GD_RT;
            hipLaunchKernelGGL(gkr_begin_eq_kernel, dim3((uint32_t)blocks), dim3(GD_RT), 0, c.stream, S, I, n_inst,
                               n_layers, L, eq);
            SR_HIP(hipGetLastError());
        }
        uint32_t rho = 0;
        for (; L - rho >= GD_SMALL_LOG; rho++) {
            const uint64_t nt = 1ull << (L - rho - 1);
            // summing workgroups per instance, plus one column for the round-start prep
            const uint64_t bx = std::min<uint64_t>((nt + GD_RT / 4 - 1) / (GD_RT / 4), GD_MAX_BX - 1) + 1;
            uint32_t n_part = 0;  // instances summed in this round: active with onv >= L - rho
            for (uint32_t k = 0; k < n_inst; k++)
                n_part += inst[k].n_vars + L >= n_layers && inst[k].n_vars + L - n_layers >= L - rho;
            KTimer kt(c, "gkr_round", round_bytes(L, rho));
            hipLaunchKernelGGL(gkr_round_kernel, dim3((uint32_t)bx, n_part), dim3(GD_RT), 0, c.stream, S, I, n_inst,
                               n_layers, L, rho, (const uint32_t*)eq, rec, run->d_part.as<unsigned long long>(),
                               run->d_cnt.as<uint32_t>());
            SR_HIP(hipGetLastError());
        }
        double bytes = 0;
        for (uint32_t r2 = rho; r2 < L; r2++) bytes += round_bytes(L, r2);
        KTimer kt(c, "gkr_small_rounds", bytes);
        hipLaunchKernelGGL(gkr_fused_kernel, dim3(1), dim3(GD_FT), 0, c.stream, S, I, n_inst, n_layers, L, L + 1, rho,
                           eq, rec);
        SR_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(gkr_chan_out_kernel, dim3(1), dim3(64), 0, c.stream, d_ch, (const GdState*)S);
    SR_HIP(hipGetLastError());
    return run.release();
}

void gkr_release(GkrRun* run) { delete run; }

// Reads the proof back (one synchronisation) and assembles the word stream; ch_out (optional)
// receives the final channel state.
static std::vector<uint32_t> gkr_collect_into(Ctx& c, GkrRun* run, Channel* ch_out) {
    std::unique_ptr<GkrRun> own(run);
    const std::vector<GkrInstance>& inst = run->inst;
    const uint32_t n_inst = run->n_inst, n_layers = run->n_layers;
    const uint32_t total_rounds = n_layers * (n_layers - 1) / 2;
    std::vector<uint32_t> hrec(run->rec_words);
    GdState* hs = (GdState*)c.pinned(sizeof(GdState));
    SR_HIP(hipMemcpyAsync(hs, run->d_state.p, sizeof(GdState), hipMemcpyDeviceToHost, c.stream));
    SR_HIP(hipMemcpyAsync(hrec.data(), run->d_rec.p, 4 * run->rec_words, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    SR_CHECK(hs->status != 1, "gkr: channel draw rejected 64 times");
    SR_CHECK(hs->status == 0, "gkr: division by zero in a round polynomial");
    // Assemble the proof word stream.
    std::vector<uint32_t> w;
    auto ncols = [&](uint32_t k) { return inst[k].kind == GKR_GRAND_PRODUCT ? 1u : 2u; };
    for (uint32_t L = 0; L < n_layers; L++) {
        w.push_back(L);
        for (uint32_t rho = 0; rho < L; rho++) {
            const uint32_t* rr = &hrec[(uint64_t)(L * (L - 1) / 2 + rho) * GD_REC_ROUND];
            SR_CHECK(rr[0] <= 4, "gkr: round record");
            w.push_back(rr[0]);
            w.insert(w.end(), rr + 1, rr + 1 + 4 * rr[0]);
        }
        uint32_t n_act = 0;
        for (uint32_t k = 0; k < n_inst; k++) n_act += inst[k].n_vars + L >= n_layers;
        w.push_back(n_act);
        for (uint32_t k = 0; k < n_inst; k++) {
            if (inst[k].n_vars + L < n_layers) continue;
            const uint32_t* rm = &hrec[(uint64_t)total_rounds * GD_REC_ROUND + ((uint64_t)L * n_inst + k) * GD_REC_MASK];
            w.push_back(k);
            w.push_back(ncols(k));
            if (ncols(k) == 1)
                w.insert(w.end(), rm + 8, rm + 16);
            else
                w.insert(w.end(), rm, rm + 16);
        }
    }
    for (uint32_t k = 0; k < n_inst; k++) {
        w.push_back(ncols(k));
        w.insert(w.end(), &hs->outputs[k][0][0], &hs->outputs[k][0][0] + 4 * ncols(k));
    }
    w.push_back(n_layers);
    w.insert(w.end(), &hs->ood[0][0], &hs->ood[0][0] + 4 * n_layers);
    for (uint32_t k = 0; k < n_inst; k++) {
        w.push_back(ncols(k));
        w.insert(w.end(), &hs->claims[k][0][0], &hs->claims[k][0][0] + 4 * ncols(k));
        w.push_back(inst[k].n_vars);
    }
    w.insert(w.end(), hs->dg, hs->dg + 8);
    if (ch_out) {
        memcpy(ch_out->digest, hs->dg, 32);
        ch_out->n_draws = hs->n_draws;
    }
    return w;
}

std::vector<uint32_t> gkr_collect(Ctx& c, GkrRun* run) { return gkr_collect_into(c, run, nullptr); }

std::vector<uint32_t> gkr_prove_batch(Ctx& c, Channel& ch, const std::vector<GkrInstance>& inst) {
    DBuf d_ch(&c, sizeof(DevChannel));
    DevChannel hc{};
    memcpy(hc.digest, ch.digest, 32);
    hc.n_draws = ch.n_draws;
    SR_HIP(hipMemcpyAsync(d_ch.p, c.stage(&hc, sizeof(hc)), sizeof(hc), hipMemcpyHostToDevice, c.stream));
    GkrRun* run = gkr_launch_on_channel(c, d_ch.as<DevChannel>(), inst);
    return gkr_collect_into(c, run, &ch);
}

}  // namespace sr
