// Wide-Fibonacci proof on one MI355X: host mirror of the reference's generic
// prover (transcript order, FRI orchestration, decommitment and proof wire)
// driving the rocm kernels.  Every heavy stage runs on the device: trace
// generation, interpolation, LDE, Merkle, composition, sampled values,
// quotients, FRI folds and PoW.  The host only keeps the Fiat-Shamir channel
// (a few hundred bytes per draw) and assembles the proof from one gather.
//
// Order of operations (src/examples/common/prover_transaction.zig:222-339,
// src/prover/prove.zig:209-389, src/prover/pcs/scheme.zig:571-800,
// src/prover/fri.zig:163-621).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <memory>
#include <string>

#include "prover.hpp"

namespace sr {

// ---------------------------------------------------------------- channel
// Blake2sChannel (src/core/channel/blake2s.zig:32-287).
void Channel::mix_bytes(const void* data, size_t n) {
    B2Stream s;
    s.update(digest, 32);
    s.update(data, n);
    s.finalize(digest);
    n_draws = 0;
}
void Channel::mix_felts(const qm31* f, size_t n) {
    std::vector<uint32_t> w(4 * n);
    for (size_t i = 0; i < n; i++) qstore(f[i], &w[4 * i]);
    mix_bytes(w.data(), 16 * n);
}
void Channel::mix_u32s(const uint32_t* d, size_t n) { mix_bytes(d, 4 * n); }
void Channel::mix_u64(uint64_t v) {
    uint32_t w[2] = {(uint32_t)v, (uint32_t)(v >> 32)};
    mix_u32s(w, 2);
}
// MerkleChannel.mixRoot (src/core/vcs_lifted/blake2_merkle.zig:237-249).
void Channel::mix_root(const uint8_t root[32]) { mix_bytes(root, 32); }
void Channel::draw_u32s(uint32_t out[8]) {
    uint8_t in[37];
    memcpy(in, digest, 32);
    memcpy(in + 32, &n_draws, 4);
    in[36] = 0;
    n_draws++;
    uint8_t h[32];
    b2_hash(in, 37, h);
    memcpy(out, h, 32);
}
void Channel::draw_base_felts(uint32_t out[8]) {
    for (;;) {
        uint32_t w[8];
        draw_u32s(w);
        bool ok = true;
        for (int i = 0; i < 8; i++)
            if (w[i] >= 2 * P) ok = false;
        if (!ok) continue;
        for (int i = 0; i < 8; i++) out[i] = mred64(w[i]);
        return;
    }
}
qm31 Channel::draw_secure_felt() {
    uint32_t f[8];
    draw_base_felts(f);
    return qload(f);
}
// Proof-of-work prefix H(LE32(0x12345678) || 0^12 || digest || LE32(bits)) (blake2s.zig:142-154).
void Channel::pow_prefix(uint32_t bits, uint8_t out[32]) const {
    uint8_t in[52];
    uint32_t pp = 0x12345678u;
    memcpy(in, &pp, 4);
    memset(in + 4, 0, 12);
    memcpy(in + 16, digest, 32);
    memcpy(in + 48, &bits, 4);
    b2_hash(in, 52, out);
}

// ---------------------------------------------------------------- helpers
static double host_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct QPoint {
    qm31 x, y;
};
static QPoint qp_double(QPoint p) {
    qm31 xx = qsqr(p.x), xy = qmul(p.x, p.y);
    return {qsub(qadd(xx, xx), qone()), qadd(xy, xy)};
}
static QPoint qp_add_m31(QPoint a, pt b) {
    qm31 bx = qfromm(b.x), by = qfromm(b.y);
    return {qsub(qmul(a.x, bx), qmul(a.y, by)), qadd(qmul(a.x, by), qmul(a.y, bx))};
}

// Witness positions of a lifted Merkle decommitment (prover/vcs_lifted/decommit.zig:171-236):
// (layer, index) of each sibling hash not derivable from the queried ones.
static void merkle_witness_positions(uint32_t max_log, const std::vector<uint64_t>& positions,
                                     std::vector<std::pair<uint32_t, uint64_t>>& out) {
    std::vector<uint64_t> prev;
    for (size_t i = 0; i < positions.size(); i++)
        if (i == 0 || positions[i - 1] != positions[i]) prev.push_back(positions[i]);
    for (int32_t layer = (int32_t)max_log - 1; layer >= 0; layer--) {
        std::vector<uint64_t> cur;
        size_t q = 0;
        while (q < prev.size()) {
            uint64_t first = prev[q];
            bool sib = q + 1 < prev.size() && (first ^ 1) == prev[q + 1];
            if (!sib) out.push_back({(uint32_t)layer + 1, first ^ 1});
            cur.push_back(first >> 1);
            q += sib ? 2 : 1;
        }
        prev.swap(cur);
    }
}


// Compact JSON writer into a growing char buffer: byte values come from a table of their
// decimal forms, wider integers are formatted two digits at a time.
struct JsonOut {
    std::string s;
    size_t n = 0;
    // 32 KiB to start (grown by doubling): blocks of 128 KiB and more come from mmap, whose
    // unmapping on free costs more than the encoding of a small proof
    JsonOut() { s.resize(1 << 15); }
    char* room(size_t k) {
        if (n + k > s.size()) s.resize(std::max(2 * s.size(), n + k));
        return &s[n];
    }
    void raw(const char* t) {
        const size_t k = strlen(t);
        memcpy(room(k), t, k);
        n += k;
    }
    void u64(uint64_t v) {
        static const char d2[] = "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
                                 "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
                                 "8081828384858687888990919293949596979899";
        char b[24];
        int k = 24;
        while (v >= 100) {
            const unsigned r = (unsigned)(v % 100);
            v /= 100;
            b[--k] = d2[2 * r + 1];
            b[--k] = d2[2 * r];
        }
        if (v >= 10) {
            b[--k] = d2[2 * v + 1];
            b[--k] = d2[2 * v];
        } else {
            b[--k] = (char)('0' + v);
        }
        memcpy(room(24 - k), b + k, 24 - k);
        n += 24 - k;
    }
    void hash(const uint8_t* h) {
        // "[b0,b1,...,b31]": at most 32 * 4 + 1 bytes
        struct Tab {
            char txt[256][4];
            uint8_t len[256];
            Tab() {
                for (int v = 0; v < 256; v++) {
                    int k = 0;
                    if (v >= 100) txt[v][k++] = (char)('0' + v / 100);
                    if (v >= 10) txt[v][k++] = (char)('0' + (v / 10) % 10);
                    txt[v][k++] = (char)('0' + v % 10);
                    len[v] = (uint8_t)k;
                }
            }
        };
        static const Tab tab;
        char* o = room(32 * 4 + 1);
        char* p = o;
        *p++ = '[';
        for (int i = 0; i < 32; i++) {
            memcpy(p, tab.txt[h[i]], 4);
            p += tab.len[h[i]];
            *p++ = ',';
        }
        p[-1] = ']';
        n += (size_t)(p - o);
    }
    void qm(const uint32_t* w) {
        raw("[");
        for (int i = 0; i < 4; i++) {
            if (i) raw(",");
            u64(w[i]);
        }
        raw("]");
    }
    std::string finish() {
        s.resize(n);
        return std::move(s);
    }
};

// ---------------------------------------------------------------- quotient batches
QuotientPlan plan_quotients(Ctx& c, uint32_t L, const std::vector<uint32_t>& col_logs,
                            const std::vector<std::vector<ColumnSample>>& samples, qm31 alpha) {
    SR_CHECK(col_logs.size() == samples.size(), "quotients: one sample list per column");
    std::vector<SamplePoint> bpts;
    std::vector<std::vector<std::pair<uint32_t, qm31>>> bent;
    std::vector<qm31> sa, sb;
    qm31 rp = qone();
    const uint32_t lift_gen = subgroup_gen(L);
    auto add_sample = [&](uint32_t col, SamplePoint p, qm31 v) {
        size_t bi = 0;
        for (; bi < bpts.size(); bi++)
            if (qeq(bpts[bi].x, p.x) && qeq(bpts[bi].y, p.y)) break;
        if (bi == bpts.size()) {
            bpts.push_back(p);
            bent.emplace_back();
            sa.push_back(qzero());
            sb.push_back(qzero());
        }
        // complexConjugateLineCoeffs (core/constraints.zig:87-97)
        SR_CHECK(!qeq(p.y, qconj(p.y)), "degenerate sample point (y in the base field)");
        qm31 a = qsub(qconj(v), v);
        qm31 cc = qsub(qconj(p.y), p.y);
        qm31 b = qsub(qmul(v, cc), qmul(a, p.y));
        bent[bi].push_back({col, qmul(rp, cc)});
        sa[bi] = qadd(sa[bi], qmul(rp, a));
        sb[bi] = qadd(sb[bi], qmul(rp, b));
        rp = qmul(rp, alpha);
    };
    for (uint32_t col = 0; col < col_logs.size(); col++) {
        const auto& ss = samples[col];
        SR_CHECK(ss.size() <= 2, "quotients: at most two samples per column");
        if (ss.size() == 2) {
            // Periodicity sample (core/pcs/quotients/samples.zig:270-285): the second point plus
            // the lifting domain's step doubled log-size times.
            const pt g = idx_to_point(c.gen, (uint32_t)(((uint64_t)lift_gen << col_logs[col]) & ORDER_MASK));
            const SamplePoint p = ss[1].p;
            const SamplePoint q = {qsub(qmul(p.x, qfromm(g.x)), qmul(p.y, qfromm(g.y))),
                                   qadd(qmul(p.x, qfromm(g.y)), qmul(p.y, qfromm(g.x)))};
            add_sample(col, q, ss[1].v);
        }
        for (const auto& smp : ss) add_sample(col, smp.p, smp.v);
    }
    QuotientPlan out;
    for (size_t bi = 0; bi < bpts.size(); bi++) {
        QuotientBatchDev bd;
        cm31 prx = bpts[bi].x.c0, pry = bpts[bi].y.c0, pix = bpts[bi].x.c1, piy = bpts[bi].y.c1;
        bd.det = csub(cmul(prx, piy), cmul(pry, pix));
        bd.pix = pix;
        bd.piy = piy;
        bd.sum_a = sa[bi];
        bd.sum_b = sb[bi];
        bd.first = (uint32_t)out.entry_cols.size();
        bd.count = (uint32_t)bent[bi].size();
        for (auto& e : bent[bi]) {
            out.entry_cols.push_back(e.first);
            uint32_t w4[4];
            qstore(e.second, w4);
            out.entry_coeffs.insert(out.entry_coeffs.end(), w4, w4 + 4);
        }
        out.batches.push_back(bd);
    }
    return out;
}

// ---------------------------------------------------------------- prove
// LDE of coefficient columns (pointer list) by `blowup` and Merkle commit
// (commit_polys.zig:10-67; preparation.zig:113-210).
void commit_coefficients(Ctx& c, TreeOwner& owner, Committed& out, const std::vector<const uint32_t*>& coeffs,
                         uint32_t coeff_log, uint32_t blowup) {
    uint32_t n = (uint32_t)coeffs.size();
    uint32_t m = coeff_log + blowup;
    uint64_t N = 1ull << m;
    out.coeffs = coeffs;
    out.coeff_logs.assign(n, coeff_log);
    out.lde_logs.assign(n, m);
    out.lde.resize(n);
    if (n) {
        out.lde_buf = DBuf(&c, (size_t)n * N * 4);
        for (uint32_t i = 0; i < n; i++) out.lde[i] = out.lde_buf.as<uint32_t>() + i * N;
        DBuf d_src = upload(c, coeffs.data(), n);
        DBuf d_dst = upload(c, out.lde.data(), n);
        launch_circle_fft(c, false, d_src.as<const uint32_t* const>(), d_dst.as<uint32_t* const>(), n, m, coeff_log);
    }
    std::vector<const uint32_t*> cols(out.lde.begin(), out.lde.end());
    out.tree = owner.add(merkle_commit(c, cols, out.lde_logs));
}

void plan_commit(Ctx& c, TreeOwner& owner, Committed& out, const std::vector<const uint32_t*>& coeffs,
                 uint32_t coeff_log, uint32_t blowup, HostPack& pack, CommitPlan& cp) {
    cp.n = (uint32_t)coeffs.size();
    cp.m = coeff_log + blowup;
    cp.coeff_log = coeff_log;
    SR_CHECK(cp.n >= 1, "commit: no columns");
    const uint64_t N = 1ull << cp.m;
    out.coeffs = coeffs;
    out.coeff_logs.assign(cp.n, coeff_log);
    out.lde_logs.assign(cp.n, cp.m);
    out.lde.resize(cp.n);
    out.lde_buf = DBuf(&c, (size_t)cp.n * N * 4);
    for (uint32_t i = 0; i < cp.n; i++) out.lde[i] = out.lde_buf.as<uint32_t>() + i * N;
    cp.at_src = pack.add(coeffs.data(), cp.n);
    cp.at_dst = pack.add(out.lde.data(), cp.n);
    // equal log sizes: the stable sort by log size (columns.zig:12-46) keeps the column order
    std::vector<const uint32_t*> cols(out.lde.begin(), out.lde.end());
    LeafPlan plan = merkle_leaf_plan(cols.data(), out.lde_logs.data(), cp.n);
    cp.at_tab = pack.add(plan.tab.data(), plan.tab.size());
    cp.segs = plan.segs;
    out.tree = owner.make(cp.m);
}

void launch_commit(Ctx& c, Committed& out, const CommitPlan& cp, uint8_t* d_pack, uint8_t* d_root_out) {
    launch_circle_fft(c, false, (const uint32_t* const*)(d_pack + cp.at_src), (uint32_t* const*)(d_pack + cp.at_dst),
                      cp.n, cp.m, cp.coeff_log);
    launch_merkle_leaves_dev(c, (const LeafColDev*)(d_pack + cp.at_tab), cp.segs, cp.n, *out.tree);
    launch_merkle_parents(c, *out.tree, nullptr, d_root_out);
}

void launch_commit_from_evals(Ctx& c, Committed& out, const CommitPlan& cp, uint8_t* d_pack, size_t at_evals,
                              uint8_t* d_root_out) {
    const uint32_t* const* ev = (const uint32_t* const*)(d_pack + at_evals);
    uint32_t* const* cf = (uint32_t* const*)(d_pack + cp.at_src);
    uint32_t* const* ld = (uint32_t* const*)(d_pack + cp.at_dst);
    if (cp.m == cp.coeff_log + 1) {
        launch_circle_ifft_lde(c, ev, cf, ld, cp.n, cp.coeff_log);
    } else {
        launch_circle_fft(c, true, ev, cf, cp.n, cp.coeff_log, cp.coeff_log);
        launch_circle_fft(c, false, (const uint32_t* const*)cf, ld, cp.n, cp.m, cp.coeff_log);
    }
    launch_merkle_leaves_dev(c, (const LeafColDev*)(d_pack + cp.at_tab), cp.segs, cp.n, *out.tree);
    launch_merkle_parents(c, *out.tree, nullptr, d_root_out);
}

void fri_fold_commit_down(Ctx& c, TreeOwner& owner, const uint32_t* d_top, uint32_t top_log, uint32_t last_log,
                          DevChannel* d_ch, uint32_t* const* dst_lines, uint32_t* d_last, uint8_t* d_roots,
                          std::vector<DeviceTree*>& trees) {
    SR_CHECK(last_log >= 1 && top_log > last_log && top_log <= 29, "FRI cascade: layer sizes");
    auto line_at = [&](uint32_t l) -> uint32_t* {
        if (l == top_log) return (uint32_t*)d_top;
        if (l == last_log) return d_last;
        return dst_lines[top_log - 1 - l];
    };
    size_t k = 0;
    // layers top_log-1 .. tail_top: one launch each (fold, commit, channel step); below, the
    // tail kernel, or a plain fold when the last layer is larger than the tail
    const uint32_t tail_top = std::max(last_log + 1, std::min(top_log, FRI_TAIL_MAX_LOG + 1));
    for (uint32_t l = top_log - 1; l >= tail_top; l--) {
        DeviceTree* t = owner.make(l);
        launch_fri_commit_line(c, line_at(l + 1), l + 1, line_at(l), *t, d_ch, d_roots + 32 * k++);
        trees.push_back(t);
    }
    if (tail_top <= FRI_TAIL_MAX_LOG + 1) {
        FriTailArgs ta{};
        ta.top_log = tail_top;
        ta.last_log = last_log;
        ta.ch = d_ch;
        ta.lines[tail_top] = (uint4*)line_at(tail_top);
        for (uint32_t l = tail_top - 1;; l--) {
            ta.lines[l] = (uint4*)line_at(l);
            if (l == last_log) break;
            DeviceTree* t = owner.make(l);
            ta.trees[l] = t->layers;
            ta.roots[l] = (uint32_t*)(d_roots + 32 * k++);
            trees.push_back(t);
        }
        launch_fri_tail(c, ta);
    } else {
        launch_fold_line(c, line_at(last_log + 1), last_log + 1, qzero(), d_last, nullptr, d_ch);
    }
}

// The proof tail in three phases, so that a batch of proofs can share the proof-of-work launch and
// the decommitment gather: the last layer (its interpolation, degree check and mix), then the
// nonce mix, the queries and the gather plan, then the wire encoding.
void finish_last_layer(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, FinishPlan& fp) {
    auto bitrev_vec = [](std::vector<qm31>& v) {
        uint32_t lg = 0;
        while ((1ull << lg) < v.size()) lg++;
        for (uint64_t i = 0; i < v.size(); i++) {
            uint64_t j = bitrev32((uint32_t)i, lg);
            if (j > i) std::swap(v[i], v[j]);
        }
    };
    const uint32_t llog = t.last_log;
    std::vector<qm31>& lv = t.last_values;
    const uint64_t lsz = lv.size();
    SR_CHECK(lsz == (1ull << llog), "last layer size");
    bitrev_vec(lv);
    {
        // line IFFT over LineDomain(halfOdds(llog)) and its doublings
        uint32_t init = half_odds_initial(llog), step = subgroup_gen(llog), clog = llog;
        while (clog > 0) {
            uint64_t csz = 1ull << clog, half = csz / 2;
            for (uint64_t s0 = 0; s0 < lsz; s0 += csz) {
                pt cur = idx_to_point(c.gen, init), stp = idx_to_point(c.gen, step);
                for (uint64_t i = 0; i < half; i++) {
                    uint32_t ix = minv(cur.x);
                    qm31 v0 = lv[s0 + i], v1 = lv[s0 + half + i];
                    lv[s0 + i] = qadd(v0, v1);
                    lv[s0 + half + i] = qmulm(qsub(v0, v1), ix);
                    cur = padd(cur, stp);
                }
            }
            init = (init * 2) & ORDER_MASK;
            step = (step * 2) & ORDER_MASK;
            clog--;
        }
        uint32_t li = minv((uint32_t)(lsz % P));
        for (auto& v : lv) v = qmulm(v, li);
    }
    bitrev_vec(lv);  // ordered coefficients
    const uint64_t bound = 1ull << prm.log_last_layer_degree_bound;
    for (uint64_t i = bound; i < lsz; i++)
        SR_CHECK(qiszero(lv[i]), "FRI last layer exceeds the degree bound (constraints not satisfied)");
    fp.lastp.assign(lv.begin(), lv.begin() + bound);
    bitrev_vec(fp.lastp);  // LinePoly.fromOrderedCoefficients
    ch.mix_felts(fp.lastp.data(), fp.lastp.size());
}

// Proof of work for one proof (scheme.zig:733-735; lowest nonce, channel/blake2s.zig:163-167).
static uint64_t grind_one(Ctx& c, const Channel& ch, uint32_t pow_bits) {
    if (pow_bits == 0) return 0;
    uint8_t prefix[32];
    ch.pow_prefix(pow_bits, prefix);
    // chunks are searched in order and each returns its lowest nonce, so the size only sets
    // how much is ground past the first solution: ~16x the expected work, 2^12..2^26
    const uint32_t cl = std::min(26u, std::max(12u, pow_bits + 4));
    const uint64_t chunk = 1ull << cl;
    for (uint64_t base = 0;; base += chunk) {
        uint64_t r = launch_pow_grind(c, prefix, pow_bits, base, chunk);
        if (r != UINT64_MAX) return r;
        SR_CHECK(base < (1ull << 50), "pow: no nonce found");
    }
}

void finish_queries(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, FinishPlan& fp, Gather& g) {
    (void)c;
    const uint32_t L = t.L;
    ch.mix_u64(fp.nonce);
    // Queries (core/queries.zig:8-60).
    std::vector<uint64_t>& qs = fp.qs;
    qs.clear();
    while (qs.size() < prm.n_queries) {
        uint32_t w[8];
        ch.draw_u32s(w);
        for (int i = 0; i < 8 && qs.size() < prm.n_queries; i++) qs.push_back(w[i] & ((1u << L) - 1));
    }
    std::sort(qs.begin(), qs.end());
    qs.erase(std::unique(qs.begin(), qs.end()), qs.end());
    const size_t nq = qs.size();

    // Decommitment plan: everything read from HBM in one gather.
    // queried values (scheme_decommit.zig; decommit.zig:77-89) and witnesses of every tree
    // with columns.
    const size_t NT = t.trees.size();
    fp.qv_at.assign(NT, 0);
    fp.tw_at.assign(NT, {});
    std::vector<std::vector<std::pair<uint32_t, uint64_t>>> tw_pos(NT);
    for (size_t tr = 0; tr < NT; tr++) {
        const Committed& T = *t.trees[tr];
        if (T.lde.empty()) continue;
        fp.qv_at[tr] = g.addr.size();
        for (size_t col = 0; col < T.lde.size(); col++)
            for (size_t q = 0; q < nq; q++) {
                uint32_t sh = T.tree->max_log - T.lde_logs[col] + 1;
                uint64_t idx = ((qs[q] >> sh) << 1) + (qs[q] & 1);
                g.add_words(T.lde[col] + idx, 1);
            }
        merkle_witness_positions(T.tree->max_log, qs, tw_pos[tr]);
        for (auto& lp : tw_pos[tr]) fp.tw_at[tr].push_back(g.add_hash(*T.tree, lp.first, lp.second));
    }
    // FRI layers (fri.zig:626-740): subsets {2k, 2k+1}, witness = non-queried partner.
    fp.lplans.clear();
    auto plan_layer = [&](const DeviceTree& lt, const std::vector<uint64_t>& q,
                          std::function<size_t(uint64_t)> add_val) {
        FinishPlan::LayerPlan lp;
        std::vector<uint64_t> dpos;
        size_t i = 0;
        while (i < q.size()) {
            uint64_t key = q[i] >> 1;
            size_t j = i + 1;
            while (j < q.size() && (q[j] >> 1) == key) j++;
            size_t at = i;
            for (uint64_t p = key << 1; p < (key << 1) + 2; p++) {
                dpos.push_back(p);
                if (at < j && q[at] == p) at++;
                else lp.wit_at.push_back(add_val(p));
            }
            i = j;
        }
        std::vector<std::pair<uint32_t, uint64_t>> hw;
        merkle_witness_positions(lt.max_log, dpos, hw);
        for (auto& h : hw) lp.hw_at.push_back(g.add_hash(lt, h.first, h.second));
        fp.lplans.push_back(lp);
    };
    plan_layer(*t.ftree0, qs, [&](uint64_t p) {
        size_t at = g.addr.size();
        for (int k = 0; k < 4; k++) g.add_words(t.qc[k] + p, 1);
        return at;
    });
    std::vector<uint64_t> lq;
    for (uint64_t q : qs)
        if (lq.empty() || lq.back() != (q >> 1)) lq.push_back(q >> 1);
    for (size_t k = 0; k < t.ftrees.size(); k++) {
        const uint32_t* base = t.lines[k];
        plan_layer(*t.ftrees[k], lq, [&](uint64_t p) { return g.add_words(base + 4 * p, 4); });
        std::vector<uint64_t> nx;
        for (uint64_t q : lq)
            if (nx.empty() || nx.back() != (q >> 1)) nx.push_back(q >> 1);
        lq.swap(nx);
    }
}

// Words of a gather plan, read from HBM in one launch through pinned host memory.
void gather_words(Ctx& c, const Gather& g, std::vector<uint32_t>& gw) {
    gw.assign(g.addr.size(), 0);
    if (g.addr.empty()) return;
    // the gather reads its address list from pinned host memory and writes the words
    // there too: no copy operations on the stream
    uint32_t* h = (uint32_t*)c.pinned(g.addr.size() * 4);
    const uint64_t* h_addr = (const uint64_t*)c.stage(g.addr.data(), g.addr.size() * 8);
    launch_gather(c, h_addr, g.addr.size(), h);
    c.sync();
    memcpy(gw.data(), h, g.addr.size() * 4);
}

std::string finish_encode(const ProveParams& prm, const ProofTail& t, const FinishPlan& fp, const uint32_t* gw) {
    const uint32_t B = prm.log_blowup_factor;
    const size_t NT = t.trees.size();
    const size_t nq = fp.qs.size();
    // Proof wire (interop/proof_wire.zig:19-76).
    JsonOut j;
    j.raw("{\"config\":{\"pow_bits\":");
    j.u64(prm.pow_bits);
    j.raw(",\"fri_config\":{\"log_blowup_factor\":");
    j.u64(B);
    j.raw(",\"log_last_layer_degree_bound\":");
    j.u64(prm.log_last_layer_degree_bound);
    j.raw(",\"n_queries\":");
    j.u64(prm.n_queries);
    j.raw(",\"fold_step\":1},\"lifting_log_size\":null},\"commitments\":[");
    for (size_t tr = 0; tr < NT; tr++) {
        if (tr) j.raw(",");
        j.hash(t.trees[tr]->tree->root);
    }
    j.raw("],\"sampled_values\":[");
    {
        size_t at = 0, col_at = 0;
        for (size_t tr = 0; tr < NT; tr++) {
            j.raw(tr ? ",[" : "[");
            for (size_t col = 0; col < t.trees[tr]->lde.size(); col++, col_at++) {
                if (col) j.raw(",");
                j.raw("[");
                for (uint32_t k = 0; k < t.n_samples[col_at]; k++) {
                    if (k) j.raw(",");
                    j.qm(&t.sv[4 * at++]);
                }
                j.raw("]");
            }
            j.raw("]");
        }
    }
    j.raw("],\"decommitments\":[");
    for (size_t tr = 0; tr < NT; tr++) {
        j.raw(tr ? ",{\"hash_witness\":[" : "{\"hash_witness\":[");
        for (size_t k = 0; k < fp.tw_at[tr].size(); k++) {
            if (k) j.raw(",");
            j.hash((const uint8_t*)&gw[fp.tw_at[tr][k]]);
        }
        j.raw("]}");
    }
    j.raw("],\"queried_values\":[");
    for (size_t tr = 0; tr < NT; tr++) {
        j.raw(tr ? ",[" : "[");
        size_t at = fp.qv_at[tr];
        for (size_t col = 0; col < t.trees[tr]->lde.size(); col++) {
            if (col) j.raw(",");
            j.raw("[");
            for (size_t q = 0; q < nq; q++) {
                if (q) j.raw(",");
                j.u64(gw[at++]);
            }
            j.raw("]");
        }
        j.raw("]");
    }
    j.raw("],\"proof_of_work\":");
    j.u64(fp.nonce);
    j.raw(",\"fri_proof\":{\"first_layer\":");
    auto emit_layer = [&](const FinishPlan::LayerPlan& lp, const DeviceTree& lt) {
        j.raw("{\"fri_witness\":[");
        for (size_t k = 0; k < lp.wit_at.size(); k++) {
            if (k) j.raw(",");
            j.qm(&gw[lp.wit_at[k]]);
        }
        j.raw("],\"decommitment\":{\"hash_witness\":[");
        for (size_t k = 0; k < lp.hw_at.size(); k++) {
            if (k) j.raw(",");
            j.hash((const uint8_t*)&gw[lp.hw_at[k]]);
        }
        j.raw("]},\"commitment\":");
        j.hash(lt.root);
        j.raw("}");
    };
    emit_layer(fp.lplans[0], *t.ftree0);
    j.raw(",\"inner_layers\":[");
    for (size_t k = 0; k < t.ftrees.size(); k++) {
        if (k) j.raw(",");
        emit_layer(fp.lplans[1 + k], *t.ftrees[k]);
    }
    j.raw("],\"last_layer_poly\":[");
    for (size_t k = 0; k < fp.lastp.size(); k++) {
        if (k) j.raw(",");
        uint32_t w4[4];
        qstore(fp.lastp[k], w4);
        j.qm(w4);
    }
    j.raw("]}}");
    return j.finish();
}

void finish_proof(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, std::string& proof,
                  const std::function<void()>& mark) {
    FinishPlan fp;
    finish_last_layer(c, ch, prm, t, fp);
    mark();  // FRI commit
    fp.nonce = grind_one(c, ch, prm.pow_bits);
    mark();  // pow
    Gather g;
    finish_queries(c, ch, prm, t, fp, g);
    std::vector<uint32_t> gw;
    gather_words(c, g, gw);
    mark();  // decommit gather
    proof = finish_encode(prm, t, fp, gw.data());
    mark();  // encode
}

// cosetVanishing(CanonicCoset(n).coset(), domain.at(0/1)) (constraints.zig:16-27): the inverse
// vanishing value at the two halves of the composition domain CanonicCoset(n + 1).
void trace_vanishing_inverses(Ctx& c, uint32_t n, uint32_t den_inv[2]) {
    const uint32_t CL = n + 1;
    for (int r = 0; r < 2; r++) {
        pt dp = idx_to_point(c.gen, domain_index_at(CL, (uint64_t)r));
        pt init = idx_to_point(c.gen, subgroup_gen(n + 1));     // odds(n).initial
        pt half_step = idx_to_point(c.gen, subgroup_gen(n + 1));  // step_size(n) / 2
        pt s = padd(padd(dp, pconj(init)), half_step);
        uint32_t x = s.x;
        for (uint32_t i = 1; i < n; i++) x = mdouble_x(x);
        den_inv[r] = minv(x);
    }
}

void prove_wide_fibonacci(Ctx& c, const ProveParams& prm, const uint32_t* d_trace_in, std::string& proof,
                          ProveStats* stats) {
    const uint32_t n = prm.log_n_rows, W = prm.sequence_len, B = prm.log_blowup_factor;
    const bool poseidon = prm.air == AIR_POSEIDON, xr = prm.air == AIR_XOR;
    SR_CHECK(prm.air == AIR_WIDE_FIBONACCI || poseidon || xr, "unknown AIR");
    SR_CHECK(n >= 1 && n <= 26, "log_n_rows out of range");
    SR_CHECK(xr || W >= 3, "sequence_len must be >= 3");
    SR_CHECK(!poseidon || W == POSEIDON_COLUMNS, "the Poseidon AIR has 1264 columns");
    SR_CHECK(!xr || (W == 1 && prm.xor_log_step <= n), "XOR AIR: one main column, log_step <= log_size");
    SR_CHECK(B >= 1 && B <= 4, "log_blowup_factor out of range");
    SR_CHECK(prm.n_queries >= 1 && prm.n_queries <= 1024, "n_queries out of range");
    SR_CHECK(prm.log_last_layer_degree_bound <= 10, "log_last_layer_degree_bound out of range");
    SR_CHECK(prm.pow_bits <= 40, "pow_bits out of range");
    const uint32_t L = n + B;  // lifting log size: every committed column is n + B
    SR_CHECK(L > prm.log_last_layer_degree_bound + B, "trace too small for the FRI configuration");
    const uint64_t N = 1ull << n;
    c.ensure_tower(L > n + 1 ? L : n + 1);

    std::vector<hipEvent_t> evs;
    std::vector<double> hms;
    // stage events only when stage timings are asked for: each event on the stream costs a
    // few microseconds of GPU time
    auto mark = [&]() {
        if (!stats) return;
        hipEvent_t e;
        SR_HIP(hipEventCreate(&e));
        SR_HIP(hipEventRecord(e, c.stream));
        evs.push_back(e);
        hms.push_back(host_ms());
    };
    mark();

    TreeOwner owner(c);
    Channel ch;
    // PcsConfig.mixInto (core/pcs/mod.zig:32-52): default fold_step, no lifting override.
    qm31 cfg = {{prm.pow_bits, B}, {prm.n_queries, prm.log_last_layer_degree_bound}};
    ch.mix_felts(&cfg, 1);

    // Tree 0: preprocessed.  Empty for wide-Fib and Poseidon (examples/wide_fibonacci/
    // trace.zig:184-186): a single leaf Blake2s(prefix) (leaves.zig:30-34), of which only the
    // root is used; for XOR the two indicator columns, committed and mixed on the device.
    Committed t0;
    if (!xr) {
        t0.tree = owner.make(0);
        uint8_t prefix[64] = {'l', 'e', 'a', 'f'};
        b2_hash(prefix, 64, t0.tree->root);
        ch.mix_root(t0.tree->root);
    }

    // From here to the FRI readback the transcript runs on a device copy of the channel:
    // the launches that need a drawn value read it from HBM, and the host reads the
    // commitments, sampled values and channel back once (transcript.hip).
    // Block `blk`: the readback region [DevChannel | roots of the first FRI layer and the inner
    // layers | last FRI line | roots of trees 1 and 2 | sampled values], the OODS point and
    // evaluation factors, then the tables of every launch up to the quotients; the host image
    // of all of it goes up in one copy before the first launch.
    const uint32_t last_log = prm.log_last_layer_degree_bound + B;
    const uint32_t n_inner = L - 1 - last_log;  // inner FRI layers log L-1 .. last_log+1
    const uint32_t NP = xr ? 2 : 0;             // preprocessed columns
    const uint32_t NS = NP + W + 8;             // sampled columns: preprocessed, trace (W), composition (8)
    const uint64_t lsz = 1ull << last_log;
    const size_t fbytes = CH_BYTES + 32 * (1 + (size_t)n_inner);
    // roots of trees 1, 2 and (XOR) 0
    const size_t at_line = fbytes, at_troots = at_line + 16 * lsz, at_sv = at_troots + 96;
    const size_t rb_bytes = at_sv + 16 * (size_t)NS;
    HostPack pack;
    {
        DevChannel hc{};
        memcpy(hc.digest, ch.digest, 32);
        hc.n_draws = ch.n_draws;
        pack.add(&hc, 1);  // offset 0
        pack.buf.resize(rb_bytes, 0);
    }
    const size_t at_trs = pack.reserve(16 * (2 + 32));  // OODS point | evaluation factors

    // Buffers of the main trace and composition phases and their tables.
    DBuf trace_own;
    const uint32_t* d_trace = d_trace_in;
    if (!d_trace) trace_own = DBuf(&c, (size_t)W * N * 4);
    if (!d_trace) d_trace = trace_own.as<uint32_t>();
    DBuf coeff1(&c, (size_t)W * N * 4);
    std::vector<const uint32_t*> cf(W);
    size_t at_ifft_src;
    {
        std::vector<const uint32_t*> src(W);
        for (uint32_t i = 0; i < W; i++) {
            src[i] = d_trace + i * N;
            cf[i] = coeff1.as<uint32_t>() + i * N;
        }
        at_ifft_src = pack.add(src.data(), W);
    }
    Committed t1;
    CommitPlan cp1;
    CommitPlan cp0;
    DBuf pre_vals, pre_coeffs;
    size_t at_pre_src = 0, at_pre_dst = 0, at_stmt = 0;
    if (xr) {
        pre_vals = DBuf(&c, 2 * N * 4);
        pre_coeffs = DBuf(&c, 2 * N * 4);
        const uint32_t* ps[2] = {pre_vals.as<uint32_t>(), pre_vals.as<uint32_t>() + N};
        std::vector<const uint32_t*> pc = {pre_coeffs.as<uint32_t>(), pre_coeffs.as<uint32_t>() + N};
        at_pre_src = pack.add(ps, 2);
        at_pre_dst = pack.add(pc.data(), 2);
        plan_commit(c, owner, t0, pc, n, B, pack, cp0);
        // mixStatement (xor.zig:599-605): mixU32s(log_size, log_step), then mixU64(offset)
        const uint32_t stw[4] = {n, prm.xor_log_step, (uint32_t)prm.xor_offset, (uint32_t)(prm.xor_offset >> 32)};
        at_stmt = pack.add(stw, 4);
    }
    plan_commit(c, owner, t1, cf, n, B, pack, cp1);

    // constraints: W - 2 recurrence steps; the Poseidon and XOR components have one
    // (poseidon.zig:414-416, xor.zig:439-441)
    const uint32_t NC = (poseidon || xr) ? 1 : W - 2, CL = n + 1;
    SR_CHECK(NC >= 1 && NC < W + 2, "constraint count");
    const uint64_t CN = 1ull << CL;
    DBuf ev_own;
    std::vector<const uint32_t*> ev(W);
    size_t at_ev_src = 0, at_ev_dst = 0;
    if (poseidon || xr) {
        // the composition evaluation reads no trace values
        for (uint32_t i = 0; i < W; i++) ev[i] = t1.lde[i];
    } else if (B == 1) {
        for (uint32_t i = 0; i < W; i++) ev[i] = t1.lde[i];
    } else {
        // trace polynomials evaluated on CanonicCoset(n + 1) from their coefficients
        ev_own = DBuf(&c, (size_t)W * CN * 4);
        std::vector<uint32_t*> dst(W);
        for (uint32_t i = 0; i < W; i++) {
            dst[i] = ev_own.as<uint32_t>() + i * CN;
            ev[i] = dst[i];
        }
        at_ev_src = pack.add(cf.data(), W);
        at_ev_dst = pack.add(dst.data(), W);
    }
    const size_t at_ev = pack.add(ev.data(), W);
    DBuf comp(&c, 4 * CN * 4);
    uint32_t* comp_c[4];
    for (int k = 0; k < 4; k++) comp_c[k] = comp.as<uint32_t>() + k * CN;
    const size_t at_cc = pack.add(comp_c, 4);
    // splitAtMid -> [left0..3, right0..3] (poly.zig:302-318; prove.zig:423-445)
    Committed t2;
    CommitPlan cp2;
    {
        std::vector<const uint32_t*> cp(8);
        for (int k = 0; k < 4; k++) {
            cp[k] = comp_c[k];
            cp[4 + k] = comp_c[k] + N;
        }
        plan_commit(c, owner, t2, cp, n, B, pack, cp2);
    }
    // Sampled values and quotients (sampled_values.zig; samples.zig:245-312): every column
    // shares log n coefficients and the lifting log.
    const Committed* trees[3] = {&t0, &t1, &t2};
    std::vector<const uint32_t*> all_coeffs;
    std::vector<uint32_t*> all_lde;
    std::vector<uint32_t> all_lde_logs;
    for (int tr = 0; tr < 3; tr++)
        for (size_t i = 0; i < trees[tr]->coeffs.size(); i++) {
            all_coeffs.push_back(trees[tr]->coeffs[i]);
            all_lde.push_back(trees[tr]->lde[i]);
            all_lde_logs.push_back(trees[tr]->lde_logs[i]);
        }
    SR_CHECK(all_coeffs.size() == NS, "sampled column count");
    const size_t at_cols = pack.add(all_coeffs.data(), NS);
    double q_in_bytes = 0;
    bool q_ident = true;  // every sampled column has the quotient domain's size
    size_t at_ents;
    {
        // the coefficients are written on the device by the quotient transcript step
        std::vector<QEntryDev> ents(NS);
        for (uint32_t i = 0; i < NS; i++) {
            SR_CHECK(all_lde_logs[i] >= 1 && all_lde_logs[i] <= L, "quotients: column larger than the domain");
            ents[i].col = all_lde[i];
            ents[i].shift = L - all_lde_logs[i] + 1;
            ents[i].pad = 0;
            q_ident = q_ident && ents[i].shift == 1;
            for (int k = 0; k < 4; k++) ents[i].coeff[k] = 0;
            q_in_bytes += 4.0 * (double)(1ull << all_lde_logs[i]);
        }
        at_ents = pack.add(ents.data(), NS);
    }
    DBuf blk = upload(c, pack.buf.data(), pack.buf.size());
    uint8_t* d_pack = blk.as<uint8_t>();
    DevChannel* d_ch = (DevChannel*)d_pack;
    uint8_t* d_roots = d_pack + CH_BYTES;
    uint32_t* d_last_line = (uint32_t*)(d_pack + at_line);
    uint8_t* d_troots = d_pack + at_troots;
    uint32_t* d_sv = (uint32_t*)(d_pack + at_sv);
    uint32_t* d_oods = (uint32_t*)(d_pack + at_trs);
    uint32_t* d_factors = d_oods + 4 * 2;

    // XOR: tree 0 from the indicator columns, mixRoot on the device.
    if (xr) {
        SR_CHECK(!d_trace_in, "XOR AIR: the trace is generated on the device");
        launch_xor_example_columns(c, n, prm.xor_log_step, prm.xor_offset, trace_own.as<uint32_t>(),
                                   pre_vals.as<uint32_t>(), pre_vals.as<uint32_t>() + N);
        launch_circle_fft(c, true, (const uint32_t* const*)(d_pack + at_pre_src),
                          (uint32_t* const*)(d_pack + at_pre_dst), 2, n, n);
        launch_commit(c, t0, cp0, d_pack, d_troots + 64);
        launch_tr_mix_draws(c, d_ch, (const uint32_t*)(d_troots + 64), 8, 0, nullptr);
    }
    // Tree 1: main trace -> coefficients -> LDE -> Merkle.
    if (!d_trace_in) {
        if (poseidon) launch_poseidon_trace(c, n, trace_own.as<uint32_t>(), N);
        else if (!xr) launch_widefib_trace(c, n, W, prm.trace_mode, prm.trace_seed, trace_own.as<uint32_t>(), N);
    }
    launch_commit_from_evals(c, t1, cp1, d_pack, at_ifft_src, d_troots);
    trace_own.reset();
    mark();  // main commit

    // mixRoot(main), mixStatement (wide_fibonacci.zig:496-498: log_n_rows, sequence_len;
    // poseidon.zig:571-573: log_n_instances), composition alpha and its powers
    // (prove.zig:237-318; wide_fibonacci/component.zig:146-228), on the device.
    DBuf d_pw(&c, 16ull * NC);
    if (xr) {
        // the drawn coefficient is unused: the one constraint's power is 1
        launch_tr_mix_draws(c, d_ch, (const uint32_t*)t1.tree->layers, 8, 0, nullptr);
        launch_tr_mix_draws(c, d_ch, (const uint32_t*)(d_pack + at_stmt), 2, 0, nullptr);
        launch_tr_mix_draws(c, d_ch, (const uint32_t*)(d_pack + at_stmt) + 2, 2, 1, d_pw.as<uint32_t>());
    } else {
        const uint32_t st[2] = {poseidon ? n + 3 : n, W};
        launch_tr_commit_draw(c, d_ch, t1.tree->layers, st, poseidon ? 1 : 2, d_pw.as<uint32_t>(), NC);
    }
    uint32_t den_inv[2];
    trace_vanishing_inverses(c, n, den_inv);
    if (poseidon || xr) {
        // compositionEval (poseidon.zig:562-569: (log_n_instances, 158, 1264, 1); xor.zig:590-597:
        // (log_size, log_step, offset, 1)) times the constraint's coefficient 1
        // (accumulation.zig:142-159)
        uint32_t cv[4] = {n + 3, POSEIDON_COLUMNS_PER_REP, POSEIDON_COLUMNS, 1};
        if (xr) {
            cv[0] = n;
            cv[1] = prm.xor_log_step;
            cv[2] = (uint32_t)(prm.xor_offset % P);
        }
        launch_constant_composition(c, CL, cv, comp_c);
    } else {
        if (B != 1)
            launch_circle_fft(c, false, (const uint32_t* const*)(d_pack + at_ev_src),
                              (uint32_t* const*)(d_pack + at_ev_dst), W, CL, n);
        launch_widefib_composition(c, (const uint32_t* const*)(d_pack + at_ev), W, CL, n, d_pw.as<uint32_t>(),
                                   den_inv[0], den_inv[1], comp_c);
    }
    // interpolate (secure_poly.zig:164-215) in place
    launch_circle_fft(c, true, (const uint32_t* const*)(d_pack + at_cc), (uint32_t* const*)(d_pack + at_cc), 4, CL, CL);
    ev_own.reset();
    launch_commit(c, t2, cp2, d_pack, d_troots + 32);
    mark();  // composition

    // mixRoot(composition), OODS point (circle.zig:159-167) and the point-evaluation
    // factors, on the device.
    launch_tr_oods(c, d_ch, t2.tree->layers, n, d_oods, d_factors);

    // Sampled values: one weight vector serves every column.
    std::vector<uint32_t> sv(4 * NS);  // read back with the FRI state
    {
        DBuf w(&c, N * 16);
        launch_eval_weights_dev(c, d_factors, n, w.as<uint32_t>());
        launch_eval_columns(c, (const uint32_t* const*)(d_pack + at_cols), NS, n, w.as<uint32_t>(), d_sv);
    }
    mark();  // sampled values

    // Quotients (samples.zig:245-312; row_evaluation.zig:73-106,402-521; constraints.zig:87-97):
    // mixFelts(sampled values), the quotient coefficient and the one batch (every sample is at
    // the OODS point: the fold count L - lde_log is 0 for every column here) on the device.
    const uint64_t LN = 1ull << L;
    DBuf quot(&c, 4 * LN * 4);
    uint32_t* qc[4];
    for (int k = 0; k < 4; k++) qc[k] = quot.as<uint32_t>() + k * LN;
    DeviceTree* ftree0 = owner.make(L);
    {
        QEntryDev* d_ents = (QEntryDev*)(d_pack + at_ents);
        DBuf d_bat(&c, sizeof(QuotientBatchDev));
        launch_tr_quotient(c, d_ch, d_sv, NS, d_oods, d_ents, d_bat.as<QuotientBatchDev>());
        // FRI (prover/fri.zig:377-621).  The commit cascade continues on the device channel: the
        // launch that produces each layer's root mixes it and draws the next alpha, which the
        // fold reads from HBM.
        launch_quotients_dev(c, d_ents, NS, d_bat.as<QuotientBatchDev>(), 1, q_in_bytes, L, qc, ftree0->layer(L), 1, 0,
                             q_ident);
        launch_merkle_parents(c, *ftree0, d_ch, d_roots);
    }
    mark();  // quotients + first layer commit

    std::vector<DBuf> lines;          // inner layer evaluations (QM31 AoS), log L-1 first
    std::vector<DeviceTree*> ftrees;  // inner layer trees
    const uint32_t* cc[4] = {qc[0], qc[1], qc[2], qc[3]};
    // the last line (log last_log) is written into the readback region
    if (L - 1 == last_log) {
        launch_fold_circle_into_line(c, cc, L, qzero(), d_last_line, false, nullptr, d_ch);
    } else {
        // layer L-1: circle fold, commit and channel step in one launch; then the cascade
        for (uint32_t k = 0; k < n_inner; k++) lines.emplace_back(&c, (1ull << (L - 1 - k)) * 16);
        DeviceTree* t = owner.make(L - 1);
        launch_fri_commit_circle(c, cc, L, lines[0].as<uint32_t>(), *t, d_ch, d_roots + 32);
        ftrees.push_back(t);
        std::vector<uint32_t*> below;
        for (uint32_t k = 1; k < n_inner; k++) below.push_back(lines[k].as<uint32_t>());
        fri_fold_commit_down(c, owner, lines[0].as<uint32_t>(), L - 1, last_log, d_ch, below.data(), d_last_line,
                             d_roots + 64, ftrees);
    }
    const uint32_t llog = last_log;
    SR_CHECK(ftrees.size() == n_inner, "FRI layer count");
    SR_CHECK(lines.size() == n_inner, "FRI line count");
    // Last layer (fri.zig:594-621; line_evaluation.zig:97-151; line.zig:134-142).
    std::vector<qm31> lv(lsz);
    {
        // one readback: channel and FRI roots, last layer, the two commitments, sampled values
        uint8_t* h = c.pinned(rb_bytes);
        SR_HIP(hipMemcpyAsync(h, d_pack, rb_bytes, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        memcpy(t1.tree->root, h + at_troots, 32);
        memcpy(t2.tree->root, h + at_troots + 32, 32);
        if (xr) memcpy(t0.tree->root, h + at_troots + 64, 32);
        memcpy(sv.data(), h + at_sv, 16 * (size_t)NS);
        DevChannel hc;
        memcpy(&hc, h, sizeof(hc));
        SR_CHECK(hc.status == 0, "FRI: channel draw rejected 64 times in a row");
        memcpy(ch.digest, hc.digest, 32);
        ch.n_draws = hc.n_draws;
        memcpy(ftree0->root, h + CH_BYTES, 32);
        for (size_t k = 0; k < ftrees.size(); k++) memcpy(ftrees[k]->root, h + CH_BYTES + 32 * (1 + k), 32);
        const uint32_t* w = (const uint32_t*)(h + at_line);
        for (uint64_t i = 0; i < lsz; i++) lv[i] = qload(w + 4 * i);
    }
    {
        ProofTail tail;
        tail.L = L;
        tail.last_log = llog;
        tail.trees = {&t0, &t1, &t2};
        tail.n_samples.assign(NS, 1);
        tail.sv = std::move(sv);
        tail.ftree0 = ftree0;
        for (int k = 0; k < 4; k++) tail.qc[k] = qc[k];
        tail.ftrees = ftrees;
        for (auto& ln : lines) tail.lines.push_back(ln.as<uint32_t>());
        tail.last_values = std::move(lv);
        finish_proof(c, ch, prm, tail, proof, mark);
    }

    if (stats) {
        SR_HIP(hipEventSynchronize(evs.back()));
        memset(stats, 0, sizeof(*stats));
        for (size_t i = 1; i < evs.size() && i < 1 + STAGES; i++) {
            float ms = 0;
            SR_HIP(hipEventElapsedTime(&ms, evs[i - 1], evs[i]));
            stats->stage_gpu_ms[i - 1] = ms;
            stats->stage_host_ms[i - 1] = hms[i] - hms[i - 1];
        }
        float tot = 0;
        SR_HIP(hipEventElapsedTime(&tot, evs.front(), evs.back()));
        stats->total_gpu_ms = tot;
        stats->total_host_ms = hms.back() - hms.front();
        stats->peak_hbm_bytes = c.peak_live;
    }
    for (auto e : evs) (void)hipEventDestroy(e);
}

// ---- host worker pool (parallel_for).  Workers are started on first use and live until the
// process ends (detached, waiting on a condition variable between jobs).  One job at a time:
// the caller that takes the pool publishes the job under the mutex, takes indices itself, and
// waits until every worker that joined the job has left it.
namespace {
struct HostPool {
    std::mutex mu;
    std::condition_variable wake, left;
    std::atomic<bool> busy{false};
    uint64_t gen = 0;  // job generation, guarded by mu
    const std::function<void(uint32_t)>* fn = nullptr;
    uint32_t n = 0, want = 0;  // indices; workers wanted by the job
    uint32_t joined = 0, inside = 0;
    std::atomic<uint32_t> next{0};
    std::exception_ptr err;
    uint32_t started = 0;

    void take(const std::function<void(uint32_t)>& f, uint32_t count) {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= count) return;
            try {
                f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(mu);
                if (!err) err = std::current_exception();
                next.store(count);
                return;
            }
        }
    }
    void worker() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            wake.wait(lk, [&] { return gen != seen && joined < want; });
            seen = gen;
            joined++;
            inside++;
            const std::function<void(uint32_t)>* f = fn;
            const uint32_t count = n;
            lk.unlock();
            take(*f, count);
            lk.lock();
            if (--inside == 0) left.notify_all();
        }
    }
};

HostPool& host_pool() {
    static HostPool* p = new HostPool();  // never destroyed: its workers outlive static destructors
    return *p;
}
}  // namespace

bool host_pool_run(uint32_t n, uint32_t threads, const std::function<void(uint32_t)>& f, std::exception_ptr& err) {
    HostPool& P = host_pool();
    bool expected = false;
    if (!P.busy.compare_exchange_strong(expected, true)) return false;
    struct Release {
        HostPool& P;
        ~Release() { P.busy.store(false); }
    } release{P};
    const uint32_t workers = threads - 1;
    {
        std::lock_guard<std::mutex> lk(P.mu);
        while (P.started < workers) {
            std::thread(&HostPool::worker, &P).detach();
            P.started++;
        }
        P.fn = &f;
        P.n = n;
        P.want = workers;
        P.joined = 0;
        P.err = nullptr;
        P.next.store(0);
        P.gen++;
    }
    P.wake.notify_all();
    P.take(f, n);
    std::unique_lock<std::mutex> lk(P.mu);
    P.want = 0;  // workers that have not joined yet stay out of this job
    P.left.wait(lk, [&] { return P.inside == 0; });
    err = P.err;
    P.fn = nullptr;
    return true;
}

}  // namespace sr
