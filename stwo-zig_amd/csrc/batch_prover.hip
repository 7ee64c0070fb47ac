// Many small wide-Fibonacci proofs in single launches.
//
// The proofs of a batch share one statement shape and PCS configuration and differ in their
// traces (seeds).  Each proof owns a slice of one device allocation, and every buffer of the
// proof sits at the same offset inside its slice: the trace, coefficients and LDE, the three
// Merkle trees, the composition, sampled values, quotients, the FRI lines and trees, and the
// device channel with its readback region.  Every stage is then one launch for the whole batch:
// the proof index is a grid dimension and a kernel reaches proof k's buffers by adding k times
// the slice size to proof 0's pointers (the FFTs take one pointer table over every proof's
// columns instead).  The transcript runs on the device channels of all proofs at once, so the
// sequence of launches up to the FRI commitment is the one-proof sequence with a batch
// dimension (prove.zig:209-389; scheme.zig:571-800; fri.zig:377-621), followed by one pitched
// readback of every proof's FRI state, one proof-of-work launch over every prefix, one gather
// of every decommitment and the wire encoding per proof.  Each proof is byte-identical to the
// one the one-proof path produces from the same parameters.
#include <string.h>

#include <chrono>

#include "prover.hpp"

namespace sr {

namespace {

double now_host_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct SliceLayout {
    size_t size = 0;
    size_t add(size_t bytes) {
        size = (size + 255) & ~(size_t)255;
        const size_t at = size;
        size += bytes;
        return at;
    }
};

}  // namespace

void prove_wide_fibonacci_batch(Ctx& c, const ProveParams& prm, uint32_t K, const uint64_t* seeds,
                                std::vector<std::string>& proofs, BatchStats* stats) {
    const double t_start = now_host_ms();
    const uint32_t n = prm.log_n_rows, W = prm.sequence_len, B = prm.log_blowup_factor;
    SR_CHECK(prm.air == AIR_WIDE_FIBONACCI, "proof batches: wide-Fibonacci AIR only");
    SR_CHECK(K >= 1 && K <= 65535, "proof batches: 1..65535 proofs");
    SR_CHECK(n >= 2 && W >= 3, "proof batches: log_n_rows >= 2 and sequence_len >= 3");
    SR_CHECK(B >= 1 && B <= 4, "log_blowup_factor out of range");
    SR_CHECK(prm.n_queries >= 1 && prm.n_queries <= 1024, "n_queries out of range");
    SR_CHECK(prm.pow_bits <= 40, "pow_bits out of range");
    SR_CHECK(prm.trace_mode <= 1, "trace_mode out of range");
    const uint32_t L = n + B;
    const uint32_t last_log = prm.log_last_layer_degree_bound + B;
    SR_CHECK(L > last_log, "trace too small for the FRI configuration");
    SR_CHECK(L - 1 <= TREE_LAYER_MAX_LOG, "proof batches: n + log_blowup must be at most 17");
    SR_CHECK(last_log <= FRI_TAIL_MAX_LOG, "proof batches: last layer above 2^10");
    const uint32_t NS = W + 8;  // sampled columns: the trace and the 8 composition halves
    SR_CHECK(NS <= 2048, "proof batches: too many columns");
    const uint64_t N = 1ull << n, LN = 1ull << L;
    const uint32_t CL = n + 1, NC = W - 2;
    const uint64_t CN = 1ull << CL;
    const uint32_t n_inner = L - 1 - last_log;
    const uint64_t lsz = 1ull << last_log;
    c.ensure_tower(L > n + 1 ? L : n + 1);

    // Transcript up to the main commitment: the PCS configuration and the empty preprocessed
    // tree's root (examples/wide_fibonacci/trace.zig:184-186), the same for every proof.
    Channel ch0;
    qm31 cfg = {{prm.pow_bits, B}, {prm.n_queries, prm.log_last_layer_degree_bound}};
    ch0.mix_felts(&cfg, 1);
    uint8_t pre_root[32];
    {
        uint8_t prefix[64] = {'l', 'e', 'a', 'f'};
        b2_hash(prefix, 64, pre_root);
    }
    ch0.mix_root(pre_root);

    // ---- slice layout.  The readback region comes first: [channel | FRI roots | last line |
    // roots of trees 1 and 2 | sampled values]; then the rest of the header (uploaded once).
    SliceLayout sl;
    const size_t at_ch = sl.add(CH_BYTES);
    const size_t at_roots = at_ch + CH_BYTES;  // first FRI layer, then the inner layers
    sl.size = at_roots + 32 * (1 + (size_t)n_inner);
    const size_t at_line = sl.add(16 * lsz);
    const size_t at_troots = sl.add(64);
    const size_t at_sv = sl.add(16 * (size_t)NS);
    const size_t rb_bytes = (sl.size + 15) & ~(size_t)15;
    const size_t at_oods = sl.add(16 * (2 + 32));
    const size_t at_ents = sl.add(sizeof(QEntryDev) * NS);
    const size_t at_bat = sl.add(sizeof(QuotientBatchDev));
    const size_t hdr_bytes = sl.size;
    const size_t at_pw = sl.add(16 * (size_t)NC);
    const size_t at_trace = sl.add(4 * (size_t)W * N);
    const size_t at_coef = sl.add(4 * (size_t)W * N);
    const size_t at_lde1 = sl.add(4 * (size_t)W * LN);
    const size_t at_ev = B != 1 ? sl.add(4 * (size_t)W * CN) : 0;
    const size_t at_comp = sl.add(16 * CN);
    const size_t at_lde2 = sl.add(32 * LN);
    const size_t tree_bytes = ((2ull << L) - 1) * 32;
    const size_t at_tree1 = sl.add(tree_bytes), at_tree2 = sl.add(tree_bytes), at_ftree0 = sl.add(tree_bytes);
    const size_t at_w = sl.add(16 * N);
    const size_t at_etab = sl.add(eval_table_bytes(n));
    const size_t at_epart = sl.add(eval_partial_bytes(NS, n));
    const size_t at_quot = sl.add(16 * LN);
    std::vector<size_t> at_lines(n_inner), at_itrees(n_inner);
    for (uint32_t k = 0; k < n_inner; k++) {
        const uint32_t l = L - 1 - k;  // layer of lines[k] and of inner tree k
        at_lines[k] = sl.add(16ull << l);
        at_itrees[k] = sl.add(((2ull << l) - 1) * 32);
    }
    const uint64_t S = ((uint64_t)sl.size + 4095) & ~(uint64_t)4095;
    SR_CHECK(S * K <= (1ull << 40), "proof batches: slab too large");

    DBuf slab(&c, S * K);
    DBuf counters(&c, 4ull * TREE_CNT_WORDS * K);
    launch_fill_u32(c, counters.as<uint32_t>(), (uint64_t)TREE_CNT_WORDS * K, 0u);
    uint8_t* base = slab.as<uint8_t>();
    uint32_t* d_cnt = counters.as<uint32_t>();
    auto P0 = [&](size_t at) { return base + at; };                // proof 0's buffer
    auto PK = [&](uint32_t k, size_t at) { return base + (uint64_t)k * S + at; };
    auto col = [&](uint32_t k, size_t at, uint64_t len, uint32_t i) { return (uint32_t*)PK(k, at) + (uint64_t)i * len; };

    // ---- host tables: pointer lists (over every proof for the FFTs, proof 0's for the rest),
    // leaf plans, the quotient entries and the initial channel
    HostPack pack;
    std::vector<const uint32_t*> t_trace;
    std::vector<uint32_t*> t_coef1w, t_lde1;
    std::vector<const uint32_t*> t_coefev;
    std::vector<uint32_t*> t_ev;
    std::vector<uint32_t*> t_comp;
    std::vector<const uint32_t*> t_cp2;
    std::vector<uint32_t*> t_lde2;
    for (uint32_t k = 0; k < K; k++) {
        for (uint32_t i = 0; i < W; i++) {
            t_trace.push_back(col(k, at_trace, N, i));
            t_coef1w.push_back(col(k, at_coef, N, i));
            t_lde1.push_back(col(k, at_lde1, LN, i));
            if (B != 1) {
                t_coefev.push_back(col(k, at_coef, N, i));
                t_ev.push_back(col(k, at_ev, CN, i));
            }
        }
        for (uint32_t q = 0; q < 4; q++) t_comp.push_back(col(k, at_comp, CN, q));
        // splitAtMid -> [left0..3, right0..3] (poly.zig:302-318; prove.zig:423-445)
        for (uint32_t h = 0; h < 2; h++)
            for (uint32_t q = 0; q < 4; q++) t_cp2.push_back(col(k, at_comp, CN, q) + h * N);
        for (uint32_t i = 0; i < 8; i++) t_lde2.push_back(col(k, at_lde2, LN, i));
    }
    const size_t tb_trace = pack.add(t_trace.data(), t_trace.size());
    const size_t tb_coef1 = pack.add(t_coef1w.data(), t_coef1w.size());
    const size_t tb_lde1 = pack.add(t_lde1.data(), t_lde1.size());
    const size_t tb_coefev = B != 1 ? pack.add(t_coefev.data(), t_coefev.size()) : 0;
    const size_t tb_ev = B != 1 ? pack.add(t_ev.data(), t_ev.size()) : 0;
    const size_t tb_comp = pack.add(t_comp.data(), t_comp.size());
    const size_t tb_cp2 = pack.add(t_cp2.data(), t_cp2.size());
    const size_t tb_lde2 = pack.add(t_lde2.data(), t_lde2.size());
    // proof 0's composition inputs (the LDE, or the evaluation on CanonicCoset(n + 1))
    std::vector<const uint32_t*> ev0(W);
    for (uint32_t i = 0; i < W; i++) ev0[i] = B == 1 ? col(0, at_lde1, LN, i) : col(0, at_ev, CN, i);
    const size_t tb_ev0 = pack.add(ev0.data(), W);
    // leaf plans of proof 0's trees 1 and 2 (one log size each: column order kept)
    std::vector<const uint32_t*> l1(W), l2(8);
    std::vector<uint32_t> l1_logs(W, L), l2_logs(8, L);
    for (uint32_t i = 0; i < W; i++) l1[i] = col(0, at_lde1, LN, i);
    for (uint32_t i = 0; i < 8; i++) l2[i] = col(0, at_lde2, LN, i);
    const LeafPlan lp1 = merkle_leaf_plan(l1.data(), l1_logs.data(), W);
    const LeafPlan lp2 = merkle_leaf_plan(l2.data(), l2_logs.data(), 8);
    const size_t tb_leaf1 = pack.add(lp1.tab.data(), lp1.tab.size());
    const size_t tb_leaf2 = pack.add(lp2.tab.data(), lp2.tab.size());
    // sampled columns of proof 0: the W trace coefficients, then the 8 composition halves
    std::vector<const uint32_t*> cols0;
    for (uint32_t i = 0; i < W; i++) cols0.push_back(col(0, at_coef, N, i));
    for (uint32_t h = 0; h < 2; h++)
        for (uint32_t q = 0; q < 4; q++) cols0.push_back(col(0, at_comp, CN, q) + h * N);
    const size_t tb_cols0 = pack.add(cols0.data(), NS);
    const uint64_t* seed_src = seeds;
    std::vector<uint64_t> seed_all;
    if (!seed_src) {
        seed_all.assign(K, prm.trace_seed);
        seed_src = seed_all.data();
    }
    const size_t tb_seeds = pack.add(seed_src, K);
    DBuf tabs = upload(c, pack.buf.data(), pack.buf.size());
    uint8_t* d_tabs = tabs.as<uint8_t>();
    auto TB = [&](size_t at) { return d_tabs + at; };

    // per-proof header: the initial channel, the quotient entries (proof 0's column pointers:
    // the kernels add the slice offset; coefficients written on the device) and a zero batch
    std::vector<uint8_t> hdr(hdr_bytes * K, 0);
    {
        std::vector<uint8_t> h(hdr_bytes, 0);
        DevChannel hc{};
        memcpy(hc.digest, ch0.digest, 32);
        hc.n_draws = ch0.n_draws;
        memcpy(h.data() + at_ch, &hc, sizeof(hc));
        // the quotient columns are the LDEs of trees 1 and 2 (lifting log L for every column:
        // shift 1), in the order of the sampled values
        QEntryDev* ents = (QEntryDev*)(h.data() + at_ents);
        for (uint32_t i = 0; i < W; i++) ents[i] = QEntryDev{col(0, at_lde1, LN, i), 1u, 0u, {0, 0, 0, 0}};
        for (uint32_t i = 0; i < 8; i++) ents[W + i] = QEntryDev{col(0, at_lde2, LN, i), 1u, 0u, {0, 0, 0, 0}};
        for (uint32_t k = 0; k < K; k++) memcpy(hdr.data() + (size_t)k * hdr_bytes, h.data(), hdr_bytes);
    }
    SR_HIP(hipMemcpy2DAsync(base, S, c.stage(hdr.data(), hdr.size()), hdr_bytes, hdr_bytes, K, hipMemcpyHostToDevice,
                            c.stream));
    DevChannel* d_ch = (DevChannel*)P0(at_ch);
    uint8_t* d_roots = P0(at_roots);

    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evs;
    SR_HIP(hipEventCreate(&evs.a));
    SR_HIP(hipEventCreate(&evs.b));
    hipEvent_t ev0e = evs.a, ev1e = evs.b;
    SR_HIP(hipEventRecord(ev0e, c.stream));

    // ---- tree 1: trace -> coefficients -> LDE -> Merkle
    launch_widefib_trace(c, n, W, prm.trace_mode, prm.trace_seed, (uint32_t*)P0(at_trace), N, K,
                         (const uint64_t*)TB(tb_seeds), S);
    if (B == 1) {
        launch_circle_ifft_lde(c, (const uint32_t* const*)TB(tb_trace), (uint32_t* const*)TB(tb_coef1),
                               (uint32_t* const*)TB(tb_lde1), K * W, n);
    } else {
        launch_circle_fft(c, true, (const uint32_t* const*)TB(tb_trace), (uint32_t* const*)TB(tb_coef1), K * W, n, n);
        launch_circle_fft(c, false, (const uint32_t* const*)TB(tb_coef1), (uint32_t* const*)TB(tb_lde1), K * W, L, n);
    }
    DeviceTree tree1{L, P0(at_tree1), {}}, tree2{L, P0(at_tree2), {}}, ftree0{L, P0(at_ftree0), {}};
    launch_merkle_leaves_dev(c, (const LeafColDev*)TB(tb_leaf1), lp1.segs, W, tree1, K, S);
    launch_merkle_parents(c, tree1, nullptr, P0(at_troots), K, S, d_cnt);

    // mixRoot(main), mixStatement (wide_fibonacci.zig:496-498), composition alpha and its powers
    const uint32_t stw[2] = {n, W};
    launch_tr_commit_draw(c, d_ch, P0(at_tree1), stw, 2, (uint32_t*)P0(at_pw), NC, K, S);
    uint32_t den_inv[2];
    trace_vanishing_inverses(c, n, den_inv);
    if (B != 1)
        launch_circle_fft(c, false, (const uint32_t* const*)TB(tb_coefev), (uint32_t* const*)TB(tb_ev), K * W, CL, n);
    uint32_t* comp0[4];
    for (uint32_t q = 0; q < 4; q++) comp0[q] = col(0, at_comp, CN, q);
    launch_widefib_composition(c, (const uint32_t* const*)TB(tb_ev0), W, CL, n, (const uint32_t*)P0(at_pw), den_inv[0],
                               den_inv[1], comp0, K, S);
    // interpolate (secure_poly.zig:164-215) in place, then tree 2
    launch_circle_fft(c, true, (const uint32_t* const*)TB(tb_comp), (uint32_t* const*)TB(tb_comp), 4 * K, CL, CL);
    launch_circle_fft(c, false, (const uint32_t* const*)TB(tb_cp2), (uint32_t* const*)TB(tb_lde2), 8 * K, L, n);
    launch_merkle_leaves_dev(c, (const LeafColDev*)TB(tb_leaf2), lp2.segs, 8, tree2, K, S);
    launch_merkle_parents(c, tree2, nullptr, P0(at_troots) + 32, K, S, d_cnt);

    // mixRoot(composition), OODS point, evaluation factors; sampled values
    uint32_t* d_oods = (uint32_t*)P0(at_oods);
    launch_tr_oods(c, d_ch, P0(at_tree2), n, d_oods, d_oods + 8, K, S);
    launch_eval_weights_dev(c, d_oods + 8, n, (uint32_t*)P0(at_w), K, S, (uint32_t*)P0(at_etab));
    launch_eval_columns_into(c, (const uint32_t* const*)TB(tb_cols0), NS, n, (const uint32_t*)P0(at_w),
                             (uint32_t*)P0(at_sv), (unsigned long long*)P0(at_epart), K, S);

    // quotients (one batch at the OODS point) with the first FRI layer's leaves, its tree and
    // channel step
    launch_tr_quotient(c, d_ch, (const uint32_t*)P0(at_sv), NS, d_oods, (QEntryDev*)P0(at_ents),
                       (QuotientBatchDev*)P0(at_bat), K, S);
    uint32_t* qc0[4];
    for (uint32_t q = 0; q < 4; q++) qc0[q] = col(0, at_quot, LN, q);
    launch_quotients_dev(c, (const QEntryDev*)P0(at_ents), NS, (const QuotientBatchDev*)P0(at_bat), 1,
                         4.0 * NS * (double)LN, L, qc0, ftree0.layer(L), K, S);
    launch_merkle_parents(c, ftree0, d_ch, d_roots, K, S, d_cnt);

    // FRI inner layers
    uint32_t* d_last = (uint32_t*)P0(at_line);
    std::vector<DeviceTree> itrees(n_inner);
    for (uint32_t k = 0; k < n_inner; k++) itrees[k] = DeviceTree{L - 1 - k, P0(at_itrees[k]), {}};
    const uint32_t* cc0[4] = {qc0[0], qc0[1], qc0[2], qc0[3]};
    if (n_inner == 0) {
        launch_fold_circle_into_line(c, cc0, L, qzero(), d_last, false, nullptr, d_ch, K, S);
    } else {
        auto line_at = [&](uint32_t l) -> uint32_t* {
            return l == last_log ? d_last : (uint32_t*)P0(at_lines[L - 1 - l]);
        };
        launch_fri_commit_circle(c, cc0, L, line_at(L - 1), itrees[0], d_ch, d_roots + 32, K, S, d_cnt);
        const uint32_t top_log = L - 1;
        size_t kr = 2;  // root slot of the next layer
        const uint32_t tail_top = std::max(last_log + 1, std::min(top_log, FRI_TAIL_MAX_LOG + 1));
        for (uint32_t l = top_log - 1; l >= tail_top; l--)
            launch_fri_commit_line(c, line_at(l + 1), l + 1, line_at(l), itrees[L - 1 - l], d_ch, d_roots + 32 * kr++,
                                   K, S, d_cnt);
        FriTailArgs ta{};
        ta.top_log = tail_top;
        ta.last_log = last_log;
        ta.ch = d_ch;
        ta.bstride = S;
        ta.lines[tail_top] = (uint4*)line_at(tail_top);
        for (uint32_t l = tail_top - 1;; l--) {
            ta.lines[l] = (uint4*)line_at(l);
            if (l == last_log) break;
            ta.trees[l] = itrees[L - 1 - l].layers;
            ta.roots[l] = (uint32_t*)(d_roots + 32 * kr++);
        }
        launch_fri_tail(c, ta, K);
        SR_CHECK(kr == 1 + (size_t)n_inner, "FRI root count");
    }

    // ---- one readback of every proof's FRI state
    const double t_enq = now_host_ms();
    uint8_t* h = c.pinned(rb_bytes * K);
    SR_HIP(hipMemcpy2DAsync(h, rb_bytes, base, S, rb_bytes, K, hipMemcpyDeviceToHost, c.stream));
    SR_HIP(hipEventRecord(ev1e, c.stream));
    c.sync();
    std::vector<uint8_t> rb(h, h + rb_bytes * K);
    float dev_ms = 0;
    SR_HIP(hipEventElapsedTime(&dev_ms, ev0e, ev1e));
    const double t_dev = now_host_ms();

    // ---- host objects of every proof: its trees and buffers at its slice
    struct PerProof {
        Committed t0, t1, t2;
        DeviceTree tr0, tr1, tr2, f0;
        std::vector<DeviceTree> fi;
        ProofTail tail;
        Channel ch;
        FinishPlan fp;
    };
    std::vector<PerProof> pp(K);
    parallel_for(K, [&](uint32_t k) {
        PerProof& P = pp[k];
        const uint8_t* r = rb.data() + (size_t)k * rb_bytes;
        DevChannel hc;
        memcpy(&hc, r + at_ch, sizeof(hc));
        SR_CHECK(hc.status == 0, "FRI: channel draw rejected 64 times in a row");
        memcpy(P.ch.digest, hc.digest, 32);
        P.ch.n_draws = hc.n_draws;
        P.tr0 = DeviceTree{0, nullptr, {}};
        memcpy(P.tr0.root, pre_root, 32);
        P.tr1 = DeviceTree{L, PK(k, at_tree1), {}};
        memcpy(P.tr1.root, r + at_troots, 32);
        P.tr2 = DeviceTree{L, PK(k, at_tree2), {}};
        memcpy(P.tr2.root, r + at_troots + 32, 32);
        P.f0 = DeviceTree{L, PK(k, at_ftree0), {}};
        memcpy(P.f0.root, r + at_roots, 32);
        P.fi.resize(n_inner);
        for (uint32_t j = 0; j < n_inner; j++) {
            P.fi[j] = DeviceTree{L - 1 - j, PK(k, at_itrees[j]), {}};
            memcpy(P.fi[j].root, r + at_roots + 32 * (1 + j), 32);
        }
        P.t0.tree = &P.tr0;
        P.t1.tree = &P.tr1;
        P.t2.tree = &P.tr2;
        for (uint32_t i = 0; i < W; i++) {
            P.t1.lde.push_back(col(k, at_lde1, LN, i));
            P.t1.lde_logs.push_back(L);
        }
        for (uint32_t i = 0; i < 8; i++) {
            P.t2.lde.push_back(col(k, at_lde2, LN, i));
            P.t2.lde_logs.push_back(L);
        }
        ProofTail& T = P.tail;
        T.L = L;
        T.last_log = last_log;
        T.trees = {&P.t0, &P.t1, &P.t2};
        T.n_samples.assign(NS, 1);
        T.sv.resize(4 * (size_t)NS);
        memcpy(T.sv.data(), r + at_sv, 16 * (size_t)NS);
        T.ftree0 = &P.f0;
        for (uint32_t q = 0; q < 4; q++) T.qc[q] = col(k, at_quot, LN, q);
        for (uint32_t j = 0; j < n_inner; j++) {
            T.ftrees.push_back(&P.fi[j]);
            T.lines.push_back((const uint32_t*)PK(k, at_lines[j]));
        }
        T.last_values.resize(lsz);
        const uint32_t* w = (const uint32_t*)(r + at_line);
        for (uint64_t i = 0; i < lsz; i++) T.last_values[i] = qload(w + 4 * i);
        finish_last_layer(c, P.ch, prm, T, P.fp);
    });

    const double t_ll = now_host_ms();
    // ---- proof of work: every prefix over one nonce range per launch, in order
    if (prm.pow_bits > 0) {
        std::vector<uint32_t> pre(8ull * K);
        for (uint32_t k = 0; k < K; k++) pp[k].ch.pow_prefix(prm.pow_bits, (uint8_t*)&pre[8ull * k]);
        DBuf d_pre = upload(c, pre.data(), pre.size());
        std::vector<uint64_t> best(K, UINT64_MAX);
        DBuf d_best = upload(c, best.data(), K);
        const uint32_t cl = std::min(26u, std::max(12u, prm.pow_bits + 4));
        const uint64_t chunk = 1ull << cl;
        for (uint64_t b0 = 0;; b0 += chunk) {
            launch_pow_grind_batch(c, d_pre.as<uint32_t>(), K, prm.pow_bits, b0, chunk,
                                   d_best.as<unsigned long long>());
            uint64_t* hb = (uint64_t*)c.pinned(8ull * K);
            SR_HIP(hipMemcpyAsync(hb, d_best.p, 8ull * K, hipMemcpyDeviceToHost, c.stream));
            c.sync();
            bool all = true;
            for (uint32_t k = 0; k < K; k++) {
                best[k] = hb[k];
                all = all && best[k] != UINT64_MAX;
            }
            if (all) break;
            SR_CHECK(b0 < (1ull << 50), "pow: no nonce found");
        }
        for (uint32_t k = 0; k < K; k++) pp[k].fp.nonce = best[k];
    }

    const double t_pow = now_host_ms();
    // ---- queries and one gather of every decommitment
    // (the plans on host threads, each into its own list; the lists concatenated for one launch)
    std::vector<Gather> gk(K);
    parallel_for(K, [&](uint32_t k) { finish_queries(c, pp[k].ch, prm, pp[k].tail, pp[k].fp, gk[k]); });
    // one pinned block: every proof's address list (written by the threads in place) and, after it,
    // the gathered words, which the encoders read where the gather left them
    std::vector<size_t> gbase(K + 1, 0);
    for (uint32_t k = 0; k < K; k++) gbase[k + 1] = gbase[k] + gk[k].addr.size();
    const size_t n_words = gbase[K];
    uint8_t* hg = c.pinned(12 * n_words + 16);
    uint64_t* h_addr = (uint64_t*)hg;
    uint32_t* h_words = (uint32_t*)(hg + 8 * n_words);
    parallel_for(K, [&](uint32_t k) {
        if (!gk[k].addr.empty()) memcpy(h_addr + gbase[k], gk[k].addr.data(), 8 * gk[k].addr.size());
    });
    if (n_words) {
        launch_gather(c, h_addr, n_words, h_words);
        c.sync();
    }
    const double t_gat = now_host_ms();
    proofs.resize(K);
    parallel_for(K, [&](uint32_t k) { proofs[k] = finish_encode(prm, pp[k].tail, pp[k].fp, h_words + gbase[k]); });
    if (stats) {
        stats->device_ms = dev_ms;
        stats->tail_ms = now_host_ms() - t_dev;
        stats->total_ms = now_host_ms() - t_start;
        stats->slice_bytes = S;
        const double t_end = now_host_ms();
        const double marks[7] = {t_start, t_enq, t_dev, t_ll, t_pow, t_gat, t_end};
        for (int i = 0; i < 6; i++) stats->phase_ms[i] = marks[i + 1] - marks[i];
    }
}

}  // namespace sr
