// GKR batch prover over device-resident layers (src/prover/lookups/
// gkr_prover.zig:34-298 with sumcheck.zig:76-197 and the oracle of
// gkr_circuit.zig:231-429).  The layer stack, the eq table, the per-round
// sums and the variable fixing run on the GPU (kernels_logup.hip); the host
// keeps the transcript, the per-instance claims and the degree-3 round
// polynomials (four QM31 coefficients each).  One device round trip per
// sum-check round serves every instance of the batch.
//
// Proof word stream (what stwo_rocm_gkr_prove_batch returns):
//   per GKR layer: n_rounds, each [len, len QM31 coefficients];
//                  n_masks, each [instance, n_columns, 2*n_columns QM31];
//   per instance: [n, output claims]; [n, ood point];
//   per instance: [n, claims to verify], n_variables; final channel digest (8 words).
#include <vector>

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

// One instance's sum-check oracle: its current layer lives in device memory.
struct Oracle {
    size_t inst;
    int kind;
    uint32_t n_vars;  // variables of the layer (the oracle has n_vars - 1)
    const uint32_t* num;
    const uint32_t* den;
    std::vector<DBuf> owned;  // buffers produced by fixing variables
    qm31 corr = qone();
};

qm31 from_acc(const uint64_t* a) { return {{mred64(a[0]), mred64(a[1])}, {mred64(a[2]), mred64(a[3])}}; }

}  // namespace

std::vector<uint32_t> gkr_prove_batch(Ctx& c, Channel& ch, const std::vector<GkrInstance>& inst) {
    SR_CHECK(!inst.empty(), "gkr: empty batch");
    const size_t n_inst = inst.size();
    uint32_t n_layers = 0;
    for (auto& in : inst) {
        SR_CHECK(in.kind >= 0 && in.kind <= 3, "gkr: layer kind");
        SR_CHECK(in.n_vars >= 1 && in.n_vars <= 30, "gkr: an instance needs 1..30 variables");
        SR_CHECK(in.d_den && (in.kind == GKR_GRAND_PRODUCT || in.kind == GKR_LOGUP_SINGLES || in.d_num),
                 "gkr: missing layer columns");
        n_layers = std::max(n_layers, in.n_vars);
    }
    // Layer stacks: layer l of instance k has l variables (buildLayers, gkr_circuit.zig:431-447).
    struct Stack {
        std::vector<DBuf> num, den;  // index l < n_vars
        std::vector<int> kind;
    };
    std::vector<Stack> st(n_inst);
    for (size_t k = 0; k < n_inst; k++) {
        const uint32_t nv = inst[k].n_vars;
        st[k].kind.assign(nv + 1, inst[k].kind == GKR_GRAND_PRODUCT ? GKR_GRAND_PRODUCT : GKR_LOGUP_GENERIC);
        st[k].kind[nv] = inst[k].kind;
        st[k].num.reserve(nv);
        st[k].den.reserve(nv);
        for (uint32_t l = 0; l < nv; l++) {
            st[k].num.emplace_back(&c, inst[k].kind == GKR_GRAND_PRODUCT ? 16 : (16ull << l));
            st[k].den.emplace_back(&c, 16ull << l);
        }
    }
    auto layer_num = [&](size_t k, uint32_t l) -> const uint32_t* {
        return l == inst[k].n_vars ? inst[k].d_num : st[k].num[l].as<uint32_t>();
    };
    auto layer_den = [&](size_t k, uint32_t l) -> const uint32_t* {
        return l == inst[k].n_vars ? inst[k].d_den : st[k].den[l].as<uint32_t>();
    };
    for (size_t k = 0; k < n_inst; k++)
        for (int32_t l = (int32_t)inst[k].n_vars - 1; l >= 0; l--)
            launch_gkr_next_layer(c, st[k].kind[l + 1], layer_num(k, l + 1), layer_den(k, l + 1), (uint32_t)l + 1,
                                  st[k].kind[l] == GKR_GRAND_PRODUCT ? nullptr : st[k].num[l].as<uint32_t>(),
                                  st[k].den[l].as<uint32_t>());
    // Output layers (gkr_circuit.zig:162-190): one download for every instance.
    std::vector<std::vector<qm31>> outputs(n_inst);
    {
        uint32_t* h = (uint32_t*)c.pinned(32 * n_inst);
        for (size_t k = 0; k < n_inst; k++) {
            const uint32_t* n0 = layer_num(k, 0);
            if (st[k].kind[0] != GKR_GRAND_PRODUCT && n0)
                SR_HIP(hipMemcpyAsync(h + 8 * k, n0, 16, hipMemcpyDeviceToHost, c.stream));
            SR_HIP(hipMemcpyAsync(h + 8 * k + 4, layer_den(k, 0), 16, hipMemcpyDeviceToHost, c.stream));
        }
        c.sync();
        for (size_t k = 0; k < n_inst; k++) {
            if (st[k].kind[0] == GKR_GRAND_PRODUCT) {
                outputs[k] = {qload(h + 8 * k + 4)};
            } else {
                qm31 nu = st[k].kind[0] == GKR_LOGUP_SINGLES ? qone() : qload(h + 8 * k);
                outputs[k] = {nu, qload(h + 8 * k + 4)};
            }
        }
    }

    std::vector<uint32_t> w;
    auto put_q = [&](qm31 v) {
        uint32_t x[4];
        qstore(v, x);
        w.insert(w.end(), x, x + 4);
    };
    std::vector<std::vector<qm31>> claims(n_inst);
    std::vector<bool> active(n_inst, false);
    std::vector<qm31> ood;
    DBuf d_acc(&c, 8 * 8 * n_inst);
    DBuf d_y(&c, 16 * (n_layers + 1));
    for (uint32_t layer = 0; layer < n_layers; layer++) {
        const uint32_t remaining = n_layers - layer;
        for (size_t k = 0; k < n_inst; k++)
            if (inst[k].n_vars == remaining) {
                claims[k] = outputs[k];
                active[k] = true;
            }
        for (size_t k = 0; k < n_inst; k++)
            if (active[k]) ch.mix_felts(claims[k].data(), claims[k].size());
        // EqEvals.generate(ood) = genEqEvals(ood[1..], 1 - ood[0]) on the device.
        const uint32_t ylen = (uint32_t)ood.size();
        DBuf d_eq(&c, ylen ? (16ull << (ylen - 1)) : 16);
        if (ylen == 0) {
            const uint32_t one[4] = {1, 0, 0, 0};
            SR_HIP(hipMemcpyAsync(d_eq.p, c.stage(one, 16), 16, hipMemcpyHostToDevice, c.stream));
        } else {
            std::vector<uint32_t> yw(4 * ylen);
            for (uint32_t i = 0; i < ylen; i++) qstore(ood[i], &yw[4 * i]);
            SR_HIP(hipMemcpyAsync(d_y.p, c.stage(yw.data(), 16 * ylen), 16 * ylen, hipMemcpyHostToDevice, c.stream));
            launch_gkr_eq_evals(c, d_y.as<uint32_t>() + 4, ylen - 1, qsub(qone(), ood[0]), d_eq.as<uint32_t>());
        }
        const qm31 sc_alpha = ch.draw_secure_felt();
        const qm31 lambda = ch.draw_secure_felt();
        // Sum-check over the active instances (sumcheck.zig:76-197).
        std::vector<Oracle> orc;
        std::vector<qm31> sc_claims;
        for (size_t k = 0; k < n_inst; k++) {
            if (!active[k]) continue;
            const uint32_t lv = inst[k].n_vars - remaining + 1;
            Oracle o;
            o.inst = k;
            o.kind = st[k].kind[lv];
            o.n_vars = lv;
            o.num = layer_num(k, lv);
            o.den = layer_den(k, lv);
            qm31 cl = qzero();  // randomLinearCombination(claims, lambda)
            for (size_t i = claims[k].size(); i > 0; i--) cl = qadd(qmul(cl, lambda), claims[k][i - 1]);
            sc_claims.push_back(cl);
            orc.push_back(std::move(o));
        }
        uint32_t sc_nv = 0;
        for (auto& o : orc) sc_nv = std::max(sc_nv, o.n_vars - 1);
        for (size_t i = 0; i < orc.size(); i++) {
            uint32_t unused = sc_nv - (orc[i].n_vars - 1);
            uint32_t p2 = 1;
            for (uint32_t u = 0; u < unused; u++) p2 = madd(p2, p2);
            sc_claims[i] = qmulm(sc_claims[i], p2);
        }
        w.push_back(sc_nv);
        std::vector<qm31> assignment;
        const uint32_t inv2 = minv(2);
        for (uint32_t round = 0; round < sc_nv; round++) {
            const uint32_t rem = sc_nv - round;
            // device sums of every instance whose variables are all still free
            launch_zero_u64(c, d_acc.as<unsigned long long>(), (uint32_t)(8 * orc.size()));
            bool any = false;
            for (size_t i = 0; i < orc.size(); i++)
                if (orc[i].n_vars - 1 == rem) {
                    launch_gkr_sums(c, orc[i].kind, d_eq.as<uint32_t>(), orc[i].num, orc[i].den, orc[i].n_vars, lambda,
                                    d_acc.as<unsigned long long>() + 8 * i);
                    any = true;
                }
            std::vector<uint64_t> acc(8 * orc.size(), 0);
            if (any) {
                uint64_t* h = (uint64_t*)c.pinned(64 * orc.size());
                SR_HIP(hipMemcpyAsync(h, d_acc.p, 64 * orc.size(), hipMemcpyDeviceToHost, c.stream));
                c.sync();
                std::copy(h, h + 8 * orc.size(), acc.begin());
            }
            std::vector<Poly> rp(orc.size());
            for (size_t i = 0; i < orc.size(); i++) {
                const uint32_t onv = orc[i].n_vars - 1;
                if (rem == onv) {
                    const qm31 f0 = qmul(from_acc(&acc[8 * i]), orc[i].corr);
                    const qm31 f2 = qmul(from_acc(&acc[8 * i + 4]), orc[i].corr);
                    rp[i] = correct_sum(f0, f2, sc_claims[i], ood, onv);
                } else {
                    rp[i].c = {qmulm(sc_claims[i], inv2)};
                    rp[i].trim();
                }
            }
            Poly comb;  // Horner combination with sc_alpha (sumcheck.zig:215-238)
            for (size_t i = orc.size(); i > 0; i--) {
                Poly s;
                if (!qiszero(sc_alpha))
                    for (auto& v : comb.c) s.c.push_back(qmul(v, sc_alpha));
                s.trim();
                Poly t;
                t.c.assign(std::max(s.c.size(), rp[i - 1].c.size()), qzero());
                for (size_t k = 0; k < t.c.size(); k++)
                    t.c[k] = qadd(k < s.c.size() ? s.c[k] : qzero(), k < rp[i - 1].c.size() ? rp[i - 1].c[k] : qzero());
                t.trim();
                comb = t;
            }
            ch.mix_felts(comb.c.data(), comb.c.size());
            const qm31 r = ch.draw_secure_felt();
            w.push_back((uint32_t)comb.c.size());
            for (auto& v : comb.c) put_q(v);
            for (size_t i = 0; i < orc.size(); i++) sc_claims[i] = rp[i].eval(r);
            for (auto& o : orc) {
                const uint32_t onv = o.n_vars - 1;
                if (rem != onv) continue;
                o.corr = qmul(o.corr, eq_one(r, ood[ood.size() - onv]));
                const uint64_t half = 1ull << (o.n_vars - 1);
                DBuf nd(&c, 16 * half);
                launch_mle_fix_first(c, false, o.den, o.n_vars, r, nd.as<uint32_t>());
                o.den = nd.as<uint32_t>();
                o.owned.push_back(std::move(nd));
                if (o.kind == GKR_LOGUP_GENERIC || o.kind == GKR_LOGUP_MULTIPLICITIES) {
                    DBuf nn(&c, 16 * half);
                    launch_mle_fix_first(c, o.kind == GKR_LOGUP_MULTIPLICITIES, o.num, o.n_vars, r,
                                         nn.as<uint32_t>());
                    o.num = nn.as<uint32_t>();
                    o.owned.push_back(std::move(nn));
                    o.kind = GKR_LOGUP_GENERIC;
                }
                o.n_vars--;
            }
            assignment.push_back(r);
        }
        // Masks of the constant oracles (gkr_circuit.zig:341-388): two values per column.
        std::vector<std::vector<std::pair<qm31, qm31>>> masks(orc.size());
        {
            uint32_t* h = (uint32_t*)c.pinned(64 * orc.size());
            for (size_t i = 0; i < orc.size(); i++) {
                const Oracle& o = orc[i];
                if (o.kind == GKR_LOGUP_GENERIC)
                    SR_HIP(hipMemcpyAsync(h + 16 * i, o.num, 32, hipMemcpyDeviceToHost, c.stream));
                else if (o.kind == GKR_LOGUP_MULTIPLICITIES)
                    SR_HIP(hipMemcpyAsync(h + 16 * i, o.num, 8, hipMemcpyDeviceToHost, c.stream));
                SR_HIP(hipMemcpyAsync(h + 16 * i + 8, o.den, 32, hipMemcpyDeviceToHost, c.stream));
            }
            c.sync();
            for (size_t i = 0; i < orc.size(); i++) {
                const Oracle& o = orc[i];
                const uint32_t* b = h + 16 * i;
                const qm31 d0 = qload(b + 8), d1 = qload(b + 12);
                if (o.kind == GKR_GRAND_PRODUCT) {
                    masks[i] = {{d0, d1}};
                } else if (o.kind == GKR_LOGUP_GENERIC) {
                    masks[i] = {{qload(b), qload(b + 4)}, {d0, d1}};
                } else if (o.kind == GKR_LOGUP_MULTIPLICITIES) {
                    masks[i] = {{qfromm(b[0]), qfromm(b[1])}, {d0, d1}};
                } else {
                    masks[i] = {{qone(), qone()}, {d0, d1}};
                }
            }
        }
        w.push_back((uint32_t)orc.size());
        for (size_t i = 0; i < orc.size(); i++) {
            std::vector<qm31> flat;
            for (auto& col : masks[i]) {
                flat.push_back(col.first);
                flat.push_back(col.second);
            }
            ch.mix_felts(flat.data(), flat.size());
            w.push_back((uint32_t)orc[i].inst);
            w.push_back((uint32_t)masks[i].size());
            for (auto& v : flat) put_q(v);
        }
        const qm31 challenge = ch.draw_secure_felt();
        ood = assignment;
        ood.push_back(challenge);
        for (size_t i = 0; i < orc.size(); i++) {
            std::vector<qm31>& cl = claims[orc[i].inst];
            cl.clear();
            for (auto& col : masks[i]) cl.push_back(qadd(qmul(challenge, qsub(col.second, col.first)), col.first));
        }
    }
    for (size_t k = 0; k < n_inst; k++) {
        w.push_back((uint32_t)outputs[k].size());
        for (auto& v : outputs[k]) put_q(v);
    }
    w.push_back((uint32_t)ood.size());
    for (auto& v : ood) put_q(v);
    for (size_t k = 0; k < n_inst; k++) {
        w.push_back(active[k] ? (uint32_t)claims[k].size() : 0u);
        if (active[k])
            for (auto& v : claims[k]) put_q(v);
        w.push_back(inst[k].n_vars);
    }
    for (int i = 0; i < 8; i++) {
        uint32_t d;
        memcpy(&d, ch.digest + 4 * i, 4);
        w.push_back(d);
    }
    return w;
}

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

}  // namespace sr
