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
#include <memory>

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

// Everything of a batch that does not depend on the traces: the slice layout, the device
// allocation, the pointer tables and leaf plans, the per-proof header image, the pinned
// readback block and (after the first graph run) the recorded launch graph.
struct BatchPlan {
    ProveParams prm{};
    uint32_t K = 0, n = 0, W = 0, B = 0, L = 0, last_log = 0, NS = 0, CL = 0, NC = 0, n_inner = 0;
    uint64_t N = 0, LN = 0, CN = 0, lsz = 0, S = 0;
    size_t at_ch = 0, at_roots = 0, at_line = 0, at_troots = 0, at_sv = 0, rb_bytes = 0, at_oods = 0, at_ents = 0,
           at_bat = 0, hdr_bytes = 0, at_pw = 0, at_trace = 0, at_coef = 0, at_lde1 = 0, at_ev = 0, at_comp = 0,
           at_lde2 = 0, at_tree1 = 0, at_tree2 = 0, at_ftree0 = 0, at_w = 0, at_etab = 0, at_epart = 0, at_quot = 0;
    std::vector<size_t> at_lines, at_itrees;
    size_t tb_trace = 0, tb_coef1 = 0, tb_lde1 = 0, tb_coefev = 0, tb_ev = 0, tb_comp = 0, tb_cp2 = 0, tb_lde2 = 0,
           tb_ev0 = 0, tb_leaf1 = 0, tb_leaf2 = 0, tb_cols0 = 0;
    Channel ch0;
    uint8_t pre_root[32] = {0};
    uint32_t den_inv[2] = {0, 0};
    LeafPlan lp1, lp2;
    DBuf slab, counters, tabs, hdr_img, seeds;
    uint8_t* h_rb = nullptr;  // pinned readback of every proof's FRI state
    bool graph_ok = false;    // no temporary buffer inside the device sequence
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the context's twiddle tower the recorded launches read: growing the tower moves it (the
    // old arrays go to the arena cache, where a trim frees them), so a replay re-records then
    const uint32_t* cap_tower[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t cap_tower_log = 0;
    Ctx* ctx = nullptr;
    ~BatchPlan() {
        drop_graph();
        if (h_rb) (void)hipHostFree(h_rb);
    }
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    bool tower_moved(const Ctx& c) const {
        return cap_tower[0] != c.d_tw || cap_tower[1] != c.d_itw || cap_tower[2] != c.d_tw2 ||
               cap_tower[3] != c.d_itw2 || cap_tower_log != c.tower_log;
    }
    uint8_t* P0(size_t at) const { return slab.as<uint8_t>() + at; }
    uint8_t* PK(uint32_t k, size_t at) const { return slab.as<uint8_t>() + (uint64_t)k * S + at; }
    uint32_t* col(uint32_t k, size_t at, uint64_t len, uint32_t i) const {
        return (uint32_t*)PK(k, at) + (uint64_t)i * len;
    }
    uint8_t* TB(size_t at) const { return tabs.as<uint8_t>() + at; }
};

BatchPlan* batch_plan_create(Ctx& c, const ProveParams& prm, uint32_t K) {
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
    std::unique_ptr<BatchPlan> bp(new BatchPlan());
    BatchPlan& P = *bp;
    P.ctx = &c;
    P.prm = prm;
    P.K = K, P.n = n, P.W = W, P.B = B, P.L = L, P.last_log = last_log, P.NS = NS;
    P.N = 1ull << n, P.LN = 1ull << L;
    P.CL = n + 1, P.NC = W - 2;
    P.CN = 1ull << P.CL;
    P.n_inner = L - 1 - last_log;
    P.lsz = 1ull << last_log;
    const uint64_t N = P.N, LN = P.LN, CN = P.CN;
    // the tables a launch of the sequence reads are built here, not in the sequence
    c.ensure_tower(L > n + 1 ? L : n + 1);
    (void)c.domain_points(L);
    (void)c.inv_circle_y(L);
    trace_vanishing_inverses(c, n, P.den_inv);
    // quotients over more than 128 columns on at most 2^16 rows use a temporary (the split form)
    P.graph_ok = !(NS > 128 && L <= 16);

    // Transcript up to the main commitment: the PCS configuration and the empty preprocessed
    // tree's root (examples/wide_fibonacci/trace.zig:184-186), the same for every proof.
    qm31 cfg = {{prm.pow_bits, B}, {prm.n_queries, prm.log_last_layer_degree_bound}};
    P.ch0.mix_felts(&cfg, 1);
    {
        uint8_t prefix[64] = {'l', 'e', 'a', 'f'};
        b2_hash(prefix, 64, P.pre_root);
    }
    P.ch0.mix_root(P.pre_root);

    // ---- slice layout.  The readback region comes first: [channel | FRI roots | last line |
    // roots of trees 1 and 2 | sampled values]; then the rest of the header.
    SliceLayout sl;
    P.at_ch = sl.add(CH_BYTES);
    P.at_roots = P.at_ch + CH_BYTES;  // first FRI layer, then the inner layers
    sl.size = P.at_roots + 32 * (1 + (size_t)P.n_inner);
    P.at_line = sl.add(16 * P.lsz);
    P.at_troots = sl.add(64);
    P.at_sv = sl.add(16 * (size_t)NS);
    P.rb_bytes = (sl.size + 15) & ~(size_t)15;
    P.at_oods = sl.add(16 * (2 + 32));
    P.at_ents = sl.add(sizeof(QEntryDev) * NS);
    P.at_bat = sl.add(sizeof(QuotientBatchDev));
    P.hdr_bytes = sl.size;
    P.at_pw = sl.add(16 * (size_t)P.NC);
    P.at_trace = sl.add(4 * (size_t)W * N);
    P.at_coef = sl.add(4 * (size_t)W * N);
    P.at_lde1 = sl.add(4 * (size_t)W * LN);
    P.at_ev = B != 1 ? sl.add(4 * (size_t)W * CN) : 0;
    P.at_comp = sl.add(16 * CN);
    P.at_lde2 = sl.add(32 * LN);
    const size_t tree_bytes = ((2ull << L) - 1) * 32;
    P.at_tree1 = sl.add(tree_bytes);
    P.at_tree2 = sl.add(tree_bytes);
    P.at_ftree0 = sl.add(tree_bytes);
    P.at_w = sl.add(16 * N);
    P.at_etab = sl.add(eval_table_bytes(n));
    P.at_epart = sl.add(eval_partial_bytes(NS, n));
    P.at_quot = sl.add(16 * LN);
    P.at_lines.resize(P.n_inner);
    P.at_itrees.resize(P.n_inner);
    for (uint32_t k = 0; k < P.n_inner; k++) {
        const uint32_t l = L - 1 - k;  // layer of lines[k] and of inner tree k
        P.at_lines[k] = sl.add(16ull << l);
        P.at_itrees[k] = sl.add(((2ull << l) - 1) * 32);
    }
    P.S = ((uint64_t)sl.size + 4095) & ~(uint64_t)4095;
    SR_CHECK(P.S * K <= (1ull << 40), "proof batches: slab too large");

    P.slab = DBuf(&c, P.S * K);
    P.counters = DBuf(&c, 4ull * TREE_CNT_WORDS * K);
    launch_fill_u32(c, P.counters.as<uint32_t>(), (uint64_t)TREE_CNT_WORDS * K, 0u);
    P.seeds = DBuf(&c, 8ull * K);

    // ---- tables: pointer lists (over every proof for the FFTs, proof 0's for the rest) and
    // leaf plans
    HostPack pack;
    std::vector<const uint32_t*> t_trace, t_coefev, t_cp2;
    std::vector<uint32_t*> t_coef1w, t_lde1, t_ev, t_comp, t_lde2;
    for (uint32_t k = 0; k < K; k++) {
        for (uint32_t i = 0; i < W; i++) {
            t_trace.push_back(P.col(k, P.at_trace, N, i));
            t_coef1w.push_back(P.col(k, P.at_coef, N, i));
            t_lde1.push_back(P.col(k, P.at_lde1, LN, i));
            if (B != 1) {
                t_coefev.push_back(P.col(k, P.at_coef, N, i));
                t_ev.push_back(P.col(k, P.at_ev, CN, i));
            }
        }
        for (uint32_t q = 0; q < 4; q++) t_comp.push_back(P.col(k, P.at_comp, CN, q));
        // splitAtMid -> [left0..3, right0..3] (poly.zig:302-318; prove.zig:423-445)
        for (uint32_t h = 0; h < 2; h++)
            for (uint32_t q = 0; q < 4; q++) t_cp2.push_back(P.col(k, P.at_comp, CN, q) + h * N);
        for (uint32_t i = 0; i < 8; i++) t_lde2.push_back(P.col(k, P.at_lde2, LN, i));
    }
    P.tb_trace = pack.add(t_trace.data(), t_trace.size());
    P.tb_coef1 = pack.add(t_coef1w.data(), t_coef1w.size());
    P.tb_lde1 = pack.add(t_lde1.data(), t_lde1.size());
    P.tb_coefev = B != 1 ? pack.add(t_coefev.data(), t_coefev.size()) : 0;
    P.tb_ev = B != 1 ? pack.add(t_ev.data(), t_ev.size()) : 0;
    P.tb_comp = pack.add(t_comp.data(), t_comp.size());
    P.tb_cp2 = pack.add(t_cp2.data(), t_cp2.size());
    P.tb_lde2 = pack.add(t_lde2.data(), t_lde2.size());
    // proof 0's composition inputs (the LDE, or the evaluation on CanonicCoset(n + 1))
    std::vector<const uint32_t*> ev0(W);
    for (uint32_t i = 0; i < W; i++) ev0[i] = B == 1 ? P.col(0, P.at_lde1, LN, i) : P.col(0, P.at_ev, CN, i);
    P.tb_ev0 = pack.add(ev0.data(), W);
    // leaf plans of proof 0's trees 1 and 2 (one log size each: column order kept)
    std::vector<const uint32_t*> l1(W), l2(8);
    std::vector<uint32_t> l1_logs(W, L), l2_logs(8, L);
    for (uint32_t i = 0; i < W; i++) l1[i] = P.col(0, P.at_lde1, LN, i);
    for (uint32_t i = 0; i < 8; i++) l2[i] = P.col(0, P.at_lde2, LN, i);
    P.lp1 = merkle_leaf_plan(l1.data(), l1_logs.data(), W);
    P.lp2 = merkle_leaf_plan(l2.data(), l2_logs.data(), 8);
    P.tb_leaf1 = pack.add(P.lp1.tab.data(), P.lp1.tab.size());
    P.tb_leaf2 = pack.add(P.lp2.tab.data(), P.lp2.tab.size());
    // sampled columns of proof 0: the W trace coefficients, then the 8 composition halves
    std::vector<const uint32_t*> cols0;
    for (uint32_t i = 0; i < W; i++) cols0.push_back(P.col(0, P.at_coef, N, i));
    for (uint32_t h = 0; h < 2; h++)
        for (uint32_t q = 0; q < 4; q++) cols0.push_back(P.col(0, P.at_comp, CN, q) + h * N);
    P.tb_cols0 = pack.add(cols0.data(), NS);
    P.tabs = upload(c, pack.buf.data(), pack.buf.size());

    // per-proof header image, copied into every slice at the start of each run: the initial
    // channel, the quotient entries (proof 0's column pointers: the kernels add the slice
    // offset; coefficients written on the device) and a zero batch
    {
        std::vector<uint8_t> h(P.hdr_bytes, 0);
        DevChannel hc{};
        memcpy(hc.digest, P.ch0.digest, 32);
        hc.n_draws = P.ch0.n_draws;
        memcpy(h.data() + P.at_ch, &hc, sizeof(hc));
        // the quotient columns are the LDEs of trees 1 and 2 (lifting log L for every column:
        // shift 1), in the order of the sampled values
        QEntryDev* ents = (QEntryDev*)(h.data() + P.at_ents);
        for (uint32_t i = 0; i < W; i++) ents[i] = QEntryDev{P.col(0, P.at_lde1, LN, i), 1u, 0u, {0, 0, 0, 0}};
        for (uint32_t i = 0; i < 8; i++) ents[W + i] = QEntryDev{P.col(0, P.at_lde2, LN, i), 1u, 0u, {0, 0, 0, 0}};
        P.hdr_img = upload(c, h.data(), h.size());
    }
    SR_HIP(hipHostMalloc((void**)&P.h_rb, P.rb_bytes * K, hipHostMallocDefault));
    c.sync();
    return bp.release();
}

// The device part of a run: header copies, every stage of every proof, and the readback into
// the plan's pinned block.  Host-side work only enqueues, so the sequence can be recorded.
static void batch_enqueue(Ctx& c, BatchPlan& P) {
    const uint32_t K = P.K, n = P.n, W = P.W, B = P.B, L = P.L, last_log = P.last_log, NS = P.NS, CL = P.CL;
    const uint64_t N = P.N, LN = P.LN, CN = P.CN, S = P.S;
    uint32_t* d_cnt = P.counters.as<uint32_t>();
    SR_HIP(hipMemcpyAsync(P.P0(0), P.hdr_img.p, P.hdr_bytes, hipMemcpyDeviceToDevice, c.stream));
    for (uint32_t k = 1; k < K; k *= 2) {
        // doubling copies of the header into slices [k, 2k)
        const uint32_t cnt = std::min(k, K - k);
        SR_HIP(hipMemcpy2DAsync(P.PK(k, 0), S, P.P0(0), S, P.hdr_bytes, cnt, hipMemcpyDeviceToDevice, c.stream));
    }
    DevChannel* d_ch = (DevChannel*)P.P0(P.at_ch);
    uint8_t* d_roots = P.P0(P.at_roots);

    // ---- tree 1: trace -> coefficients -> LDE -> Merkle
    launch_widefib_trace(c, n, W, P.prm.trace_mode, P.prm.trace_seed, (uint32_t*)P.P0(P.at_trace), N, K,
                         P.seeds.as<const uint64_t>(), S);
    if (B == 1) {
        launch_circle_ifft_lde(c, (const uint32_t* const*)P.TB(P.tb_trace), (uint32_t* const*)P.TB(P.tb_coef1),
                               (uint32_t* const*)P.TB(P.tb_lde1), K * W, n);
    } else {
        launch_circle_fft(c, true, (const uint32_t* const*)P.TB(P.tb_trace), (uint32_t* const*)P.TB(P.tb_coef1), K * W,
                          n, n);
        launch_circle_fft(c, false, (const uint32_t* const*)P.TB(P.tb_coef1), (uint32_t* const*)P.TB(P.tb_lde1), K * W,
                          L, n);
    }
    DeviceTree tree1{L, P.P0(P.at_tree1), {}}, tree2{L, P.P0(P.at_tree2), {}}, ftree0{L, P.P0(P.at_ftree0), {}};
    launch_merkle_leaves_dev(c, (const LeafColDev*)P.TB(P.tb_leaf1), P.lp1.segs, W, tree1, K, S);
    launch_merkle_parents(c, tree1, nullptr, P.P0(P.at_troots), K, S, d_cnt);

    // mixRoot(main), mixStatement (wide_fibonacci.zig:496-498), composition alpha and its powers
    const uint32_t stw[2] = {n, W};
    launch_tr_commit_draw(c, d_ch, P.P0(P.at_tree1), stw, 2, (uint32_t*)P.P0(P.at_pw), P.NC, K, S);
    if (B != 1)
        launch_circle_fft(c, false, (const uint32_t* const*)P.TB(P.tb_coefev), (uint32_t* const*)P.TB(P.tb_ev), K * W,
                          CL, n);
    uint32_t* comp0[4];
    for (uint32_t q = 0; q < 4; q++) comp0[q] = P.col(0, P.at_comp, CN, q);
    launch_widefib_composition(c, (const uint32_t* const*)P.TB(P.tb_ev0), W, CL, n, (const uint32_t*)P.P0(P.at_pw),
                               P.den_inv[0], P.den_inv[1], comp0, K, S);
    // interpolate (secure_poly.zig:164-215) in place, then tree 2
    launch_circle_fft(c, true, (const uint32_t* const*)P.TB(P.tb_comp), (uint32_t* const*)P.TB(P.tb_comp), 4 * K, CL,
                      CL);
    launch_circle_fft(c, false, (const uint32_t* const*)P.TB(P.tb_cp2), (uint32_t* const*)P.TB(P.tb_lde2), 8 * K, L,
                      n);
    launch_merkle_leaves_dev(c, (const LeafColDev*)P.TB(P.tb_leaf2), P.lp2.segs, 8, tree2, K, S);
    launch_merkle_parents(c, tree2, nullptr, P.P0(P.at_troots) + 32, K, S, d_cnt);

    // mixRoot(composition), OODS point, evaluation factors; sampled values
    uint32_t* d_oods = (uint32_t*)P.P0(P.at_oods);
    launch_tr_oods(c, d_ch, P.P0(P.at_tree2), n, d_oods, d_oods + 8, K, S);
    launch_eval_weights_dev(c, d_oods + 8, n, (uint32_t*)P.P0(P.at_w), K, S, (uint32_t*)P.P0(P.at_etab));
    launch_eval_columns_into(c, (const uint32_t* const*)P.TB(P.tb_cols0), NS, n, (const uint32_t*)P.P0(P.at_w),
                             (uint32_t*)P.P0(P.at_sv), (unsigned long long*)P.P0(P.at_epart), K, S);

    // quotients (one batch at the OODS point) with the first FRI layer's leaves, its tree and
    // channel step
    launch_tr_quotient(c, d_ch, (const uint32_t*)P.P0(P.at_sv), NS, d_oods, (QEntryDev*)P.P0(P.at_ents),
                       (QuotientBatchDev*)P.P0(P.at_bat), K, S);
    uint32_t* qc0[4];
    for (uint32_t q = 0; q < 4; q++) qc0[q] = P.col(0, P.at_quot, LN, q);
    launch_quotients_dev(c, (const QEntryDev*)P.P0(P.at_ents), NS, (const QuotientBatchDev*)P.P0(P.at_bat), 1,
                         4.0 * NS * (double)LN, L, qc0, ftree0.layer(L), K, S, true);
    launch_merkle_parents(c, ftree0, d_ch, d_roots, K, S, d_cnt);

    // FRI inner layers
    uint32_t* d_last = (uint32_t*)P.P0(P.at_line);
    std::vector<DeviceTree> itrees(P.n_inner);
    for (uint32_t k = 0; k < P.n_inner; k++) itrees[k] = DeviceTree{L - 1 - k, P.P0(P.at_itrees[k]), {}};
    const uint32_t* cc0[4] = {qc0[0], qc0[1], qc0[2], qc0[3]};
    if (P.n_inner == 0) {
        launch_fold_circle_into_line(c, cc0, L, qzero(), d_last, false, nullptr, d_ch, K, S);
    } else {
        auto line_at = [&](uint32_t l) -> uint32_t* {
            return l == last_log ? d_last : (uint32_t*)P.P0(P.at_lines[L - 1 - l]);
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
        SR_CHECK(kr == 1 + (size_t)P.n_inner, "FRI root count");
    }
    // ---- one readback of every proof's FRI state
    SR_HIP(hipMemcpy2DAsync(P.h_rb, P.rb_bytes, P.P0(0), S, P.rb_bytes, K, hipMemcpyDeviceToHost, c.stream));
}

void batch_plan_run(Ctx& c, BatchPlan& P, const uint64_t* seeds, bool use_graph, std::vector<std::string>& proofs,
                    BatchStats* stats) {
    const double t_start = now_host_ms();
    SR_CHECK(P.ctx == &c, "proof batches: the plan belongs to another context");
    const uint32_t K = P.K, W = P.W, L = P.L, last_log = P.last_log, NS = P.NS, n_inner = P.n_inner;
    const uint64_t LN = P.LN, lsz = P.lsz;
    const ProveParams& prm = P.prm;
    {
        std::vector<uint64_t> sd(K, prm.trace_seed);
        if (seeds) std::copy(seeds, seeds + K, sd.begin());
        SR_HIP(hipMemcpyAsync(P.seeds.p, c.stage(sd.data(), 8ull * K), 8ull * K, hipMemcpyHostToDevice, c.stream));
    }
    struct Events {
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } evs;
    SR_HIP(hipEventCreate(&evs.a));
    SR_HIP(hipEventCreate(&evs.b));
    SR_HIP(hipEventRecord(evs.a, c.stream));
    // a graph replays the recorded launches; the stage profile needs its own events, so a
    // profiled context runs the sequence eagerly
    const bool graph = use_graph && P.graph_ok && !c.profiling;
    if (graph && P.exec && P.tower_moved(c)) {
        SR_HIP(hipStreamSynchronize(c.stream));  // an earlier replay may still run
        P.drop_graph();
    }
    if (graph && !P.exec) {
        SR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
        try {
            batch_enqueue(c, P);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(c.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SR_HIP(hipStreamEndCapture(c.stream, &P.graph));
        SR_HIP(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
        P.cap_tower[0] = c.d_tw;
        P.cap_tower[1] = c.d_itw;
        P.cap_tower[2] = c.d_tw2;
        P.cap_tower[3] = c.d_itw2;
        P.cap_tower_log = c.tower_log;
    }
    if (graph) SR_HIP(hipGraphLaunch(P.exec, c.stream));
    else batch_enqueue(c, P);
    SR_HIP(hipEventRecord(evs.b, c.stream));
    const double t_enq = now_host_ms();
    c.sync();
    std::vector<uint8_t> rb(P.h_rb, P.h_rb + P.rb_bytes * K);
    float dev_ms = 0;
    SR_HIP(hipEventElapsedTime(&dev_ms, evs.a, evs.b));
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
        PerProof& Q = pp[k];
        const uint8_t* r = rb.data() + (size_t)k * P.rb_bytes;
        DevChannel hc;
        memcpy(&hc, r + P.at_ch, sizeof(hc));
        SR_CHECK(hc.status == 0, "FRI: channel draw rejected 64 times in a row");
        memcpy(Q.ch.digest, hc.digest, 32);
        Q.ch.n_draws = hc.n_draws;
        Q.tr0 = DeviceTree{0, nullptr, {}};
        memcpy(Q.tr0.root, P.pre_root, 32);
        Q.tr1 = DeviceTree{L, P.PK(k, P.at_tree1), {}};
        memcpy(Q.tr1.root, r + P.at_troots, 32);
        Q.tr2 = DeviceTree{L, P.PK(k, P.at_tree2), {}};
        memcpy(Q.tr2.root, r + P.at_troots + 32, 32);
        Q.f0 = DeviceTree{L, P.PK(k, P.at_ftree0), {}};
        memcpy(Q.f0.root, r + P.at_roots, 32);
        Q.fi.resize(n_inner);
        for (uint32_t j = 0; j < n_inner; j++) {
            Q.fi[j] = DeviceTree{L - 1 - j, P.PK(k, P.at_itrees[j]), {}};
            memcpy(Q.fi[j].root, r + P.at_roots + 32 * (1 + j), 32);
        }
        Q.t0.tree = &Q.tr0;
        Q.t1.tree = &Q.tr1;
        Q.t2.tree = &Q.tr2;
        for (uint32_t i = 0; i < W; i++) {
            Q.t1.lde.push_back(P.col(k, P.at_lde1, LN, i));
            Q.t1.lde_logs.push_back(L);
        }
        for (uint32_t i = 0; i < 8; i++) {
            Q.t2.lde.push_back(P.col(k, P.at_lde2, LN, i));
            Q.t2.lde_logs.push_back(L);
        }
        ProofTail& T = Q.tail;
        T.L = L;
        T.last_log = last_log;
        T.trees = {&Q.t0, &Q.t1, &Q.t2};
        T.n_samples.assign(NS, 1);
        T.sv.resize(4 * (size_t)NS);
        memcpy(T.sv.data(), r + P.at_sv, 16 * (size_t)NS);
        T.ftree0 = &Q.f0;
        for (uint32_t q = 0; q < 4; q++) T.qc[q] = P.col(k, P.at_quot, LN, q);
        for (uint32_t j = 0; j < n_inner; j++) {
            T.ftrees.push_back(&Q.fi[j]);
            T.lines.push_back((const uint32_t*)P.PK(k, P.at_lines[j]));
        }
        T.last_values.resize(lsz);
        const uint32_t* w = (const uint32_t*)(r + P.at_line);
        for (uint64_t i = 0; i < lsz; i++) T.last_values[i] = qload(w + 4 * i);
        finish_last_layer(c, Q.ch, prm, T, Q.fp);
    });

    const double t_ll = now_host_ms();
    // ---- proof of work: every prefix over one nonce range per launch, in order
    if (prm.pow_bits > 0) {
        std::vector<uint32_t> pre(8ull * K);
        for (uint32_t k = 0; k < K; k++) pp[k].ch.pow_prefix(prm.pow_bits, (uint8_t*)&pre[8ull * k]);
        DBuf d_pre = upload(c, pre.data(), pre.size());
        std::vector<uint64_t> best(K, UINT64_MAX);
        DBuf d_best = upload(c, best.data(), K);
        // 2^(bits+3) nonces per proof and launch: a proof misses its range with probability
        // e^-8, and only the proofs still without a nonce grind the next range
        const uint32_t cl = std::min(26u, std::max(12u, prm.pow_bits + 3));
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
    parallel_for(K, [&](uint32_t k) {
        proofs[k] = finish_encode(prm, pp[k].tail, pp[k].fp, h_words + gbase[k]);
        // the proof's host objects (tens of small vectors) are freed here, on the pool, rather
        // than one after another when the vectors go out of scope
        pp[k] = PerProof();
        gk[k] = Gather();
    });
    if (stats) {
        stats->device_ms = dev_ms;
        stats->tail_ms = now_host_ms() - t_dev;
        stats->total_ms = now_host_ms() - t_start;
        stats->slice_bytes = P.S;
        const double t_end = now_host_ms();
        const double marks[7] = {t_start, t_enq, t_dev, t_ll, t_pow, t_gat, t_end};
        for (int i = 0; i < 6; i++) stats->phase_ms[i] = marks[i + 1] - marks[i];
    }
}

void batch_plan_free(BatchPlan* P) { delete P; }

uint64_t batch_plan_slice_bytes(const BatchPlan& P) { return P.S; }

void prove_wide_fibonacci_batch(Ctx& c, const ProveParams& prm, uint32_t K, const uint64_t* seeds,
                                std::vector<std::string>& proofs, BatchStats* stats) {
    std::unique_ptr<BatchPlan> P(batch_plan_create(c, prm, K));
    batch_plan_run(c, *P, seeds, false, proofs, stats);
}

}  // namespace sr
