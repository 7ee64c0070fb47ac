// LogUp workload on one GPU: the XOR-lookup AIR from trace to quotient commitment.
//
// Steps (every array device-resident, all columns on the trace domain of log n >= 16):
//   1. trace a, b, c (a, b uniform 8-bit, c = a ^ b) and the table multiplicities m
//      (2^16 entries, zero-padded to n rows);
//   2. channel <- (n, seed); commit [a, b, c, m] (interpolate, LDE at blowup 2, lifted
//      Merkle), mix the root (commit_polys.zig:10-67);
//   3. z, alpha <- channel; denominators z - (a + b alpha + c alpha^2) for lookups and the
//      table rows (x, y, x ^ y);
//   4. cumulative columns S_lookup = prefix sums of 1/d and S_table = prefix sums of -m/d
//      (logup.zig:74-91); claimed sums mixed; their sum must be zero;
//   5. commit the 8 coordinate columns of S_lookup and S_table, mix the root;
//   6. batch GKR proof of the two fraction sums (gkr_prover.zig:34-298) on the channel;
//   7. OODS point (circle.zig:159-167), sampled values of the 12 committed polynomials
//      (point_evaluation.zig:14-38), mixed;
//   8. quotient accumulation of the 12 columns at the OODS point (row_evaluation.zig:402-521)
//      and its commitment (4 coordinate columns, fused leaf hashing), root mixed.
// Output word stream: main root (8), interaction root (8), quotient root (8), claimed sums
// (2 x 4), sampled values (12 x 4), GKR word count and words, final channel digest (8).
#include <chrono>
#include <vector>

#include "prover.hpp"

namespace sr {

static double lp_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

std::vector<uint32_t> logup_xor_pipeline(Ctx& c, uint32_t log_n, uint64_t seed, double* total_ms) {
    SR_CHECK(log_n >= 16 && log_n <= 26, "logup workload: log_n must be in 16..26");
    const double t0 = lp_ms();
    const uint32_t n = log_n, B = 1, L = n + B;
    const uint64_t N = 1ull << n;
    c.ensure_tower(L + 1);
    TreeOwner owner(c);
    Channel ch;
    {
        uint32_t st[3] = {n, (uint32_t)seed, (uint32_t)(seed >> 32)};
        ch.mix_u32s(st, 3);
    }
    // 1. trace: columns a, b, c, m (evaluations in storage order) then coefficients in place
    DBuf main(&c, 4 * N * 4);
    uint32_t* col[4];
    for (int k = 0; k < 4; k++) col[k] = main.as<uint32_t>() + k * N;
    DBuf mult16(&c, 4u << 16), negm(&c, 4 * N);
    launch_xor_trace(c, n, seed, col[0], col[1], col[2], mult16.as<uint32_t>());
    launch_logup_table_cols(c, mult16.as<uint32_t>(), N, col[3], negm.as<uint32_t>());
    // evaluations are kept for the denominators; coefficients go to a second buffer
    DBuf main_cf(&c, 4 * N * 4);
    std::vector<const uint32_t*> mcf(4);
    {
        std::vector<uint32_t*> dst(4);
        for (int k = 0; k < 4; k++) dst[k] = main_cf.as<uint32_t>() + k * N;
        DBuf d_src = upload(c, col, 4), d_dst = upload(c, dst.data(), 4);
        launch_circle_fft(c, true, d_src.as<const uint32_t* const>(), d_dst.as<uint32_t* const>(), 4, n, n);
        for (int k = 0; k < 4; k++) mcf[k] = dst[k];
    }
    Committed t_main;
    commit_coefficients(c, owner, t_main, mcf, n, B);
    tree_read_root(c, *t_main.tree);
    ch.mix_root(t_main.tree->root);

    // 3. relation challenges and denominators
    const qm31 z = ch.draw_secure_felt();
    const qm31 alpha = ch.draw_secure_felt();
    DBuf dl(&c, 16 * N), dt(&c, 16 * N);
    launch_logup_denominators(c, col[0], col[1], col[2], N, z, alpha, dl.as<uint32_t>());
    launch_logup_denominators(c, nullptr, nullptr, nullptr, N, z, alpha, dt.as<uint32_t>());

    // 4. cumulative columns
    DBuf sl(&c, 16 * N), stt(&c, 16 * N);
    const qm31 cl = logup_cumulative(c, 2, nullptr, dl.as<uint32_t>(), N, sl.as<uint32_t>());
    const qm31 ct = logup_cumulative(c, 1, negm.as<uint32_t>(), dt.as<uint32_t>(), N, stt.as<uint32_t>());
    SR_CHECK(qiszero(qadd(cl, ct)), "logup: lookups and table do not balance");
    {
        qm31 cs[2] = {cl, ct};
        ch.mix_felts(cs, 2);
    }
    // 5. interaction commitment: 8 coordinate columns
    DBuf inter(&c, 8 * N * 4);
    std::vector<const uint32_t*> icf(8);
    {
        uint32_t* ic[8];
        for (int k = 0; k < 8; k++) ic[k] = inter.as<uint32_t>() + k * N;
        launch_qm31_to_coords(c, sl.as<uint32_t>(), N, ic);
        launch_qm31_to_coords(c, stt.as<uint32_t>(), N, ic + 4);
        DBuf d_cc = upload(c, ic, 8);
        launch_circle_fft(c, true, d_cc.as<const uint32_t* const>(), d_cc.as<uint32_t* const>(), 8, n, n);
        for (int k = 0; k < 8; k++) icf[k] = ic[k];
    }
    Committed t_int;
    commit_coefficients(c, owner, t_int, icf, n, B);
    tree_read_root(c, *t_int.tree);
    ch.mix_root(t_int.tree->root);

    // 6. GKR batch proof of the two fraction sums
    std::vector<uint32_t> gkr = gkr_prove_batch(
        c, ch,
        {GkrInstance{GKR_LOGUP_SINGLES, nullptr, dl.as<uint32_t>(), n},
         GkrInstance{GKR_LOGUP_MULTIPLICITIES, negm.as<uint32_t>(), dt.as<uint32_t>(), n}});

    // 7. OODS point and sampled values of the 12 polynomials (one weight vector)
    const qm31 tt = ch.draw_secure_felt();
    const qm31 t2 = qsqr(tt);
    const qm31 inv1 = qinv(qadd(t2, qone()));
    const qm31 px = qmul(qsub(qone(), t2), inv1), py = qmul(qadd(tt, tt), inv1);
    std::vector<const uint32_t*> all_cf(mcf);
    all_cf.insert(all_cf.end(), icf.begin(), icf.end());
    std::vector<uint32_t*> all_lde(t_main.lde);
    all_lde.insert(all_lde.end(), t_int.lde.begin(), t_int.lde.end());
    const uint32_t NS = (uint32_t)all_cf.size();
    std::vector<uint32_t> sv(4 * NS);
    {
        qm31 f[32];
        f[0] = py;
        f[1] = px;
        for (uint32_t k = 2; k < n; k++) {
            qm31 s2 = qsqr(f[k - 1]);
            f[k] = qsub(qadd(s2, s2), qone());
        }
        DBuf w(&c, N * 16);
        launch_eval_weights(c, f, n, w.as<uint32_t>());
        DBuf d_cols = upload(c, all_cf.data(), NS);
        DBuf d_out(&c, (size_t)NS * 16);
        launch_eval_columns(c, d_cols.as<const uint32_t* const>(), NS, n, w.as<uint32_t>(), d_out.as<uint32_t>());
        uint32_t* h = (uint32_t*)c.pinned(NS * 16);
        SR_HIP(hipMemcpyAsync(h, d_out.p, NS * 16, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        memcpy(sv.data(), h, NS * 16);
    }
    ch.mix_bytes(sv.data(), sv.size() * 4);

    // 8. quotients at the OODS point, one batch, powers of the quotient alpha in column order
    const qm31 qa = ch.draw_secure_felt();
    QuotientBatchDev bd;
    std::vector<uint32_t> ecols(NS), ecf(4 * NS);
    {
        qm31 rp = qone(), sa = qzero(), sb = qzero();
        const qm31 cc = qsub(qconj(py), py);
        for (uint32_t i = 0; i < NS; i++) {
            const qm31 v = qload(&sv[4 * i]);
            const qm31 a = qsub(qconj(v), v);
            const qm31 b = qsub(qmul(v, cc), qmul(a, py));
            ecols[i] = i;
            qstore(qmul(rp, cc), &ecf[4 * i]);
            sa = qadd(sa, qmul(rp, a));
            sb = qadd(sb, qmul(rp, b));
            rp = qmul(rp, qa);
        }
        bd.det = csub(cmul(px.c0, py.c1), cmul(py.c0, px.c1));
        bd.pix = px.c1;
        bd.piy = py.c1;
        bd.sum_a = sa;
        bd.sum_b = sb;
        bd.first = 0;
        bd.count = NS;
    }
    const uint64_t LN = 1ull << L;
    DBuf quot(&c, 4 * LN * 4);
    uint32_t* qc[4];
    for (int k = 0; k < 4; k++) qc[k] = quot.as<uint32_t>() + k * LN;
    DeviceTree* qtree = owner.make(L);
    {
        std::vector<uint32_t> logs(NS, L);
        launch_quotients(c, all_lde.data(), logs.data(), ecols.data(), ecf.data(), NS, &bd, 1, L, qc,
                         qtree->layer(L));
        launch_merkle_parents(c, *qtree);
    }
    tree_read_root(c, *qtree);
    ch.mix_root(qtree->root);

    std::vector<uint32_t> out;
    auto put_root = [&](const uint8_t* r) {
        for (int i = 0; i < 8; i++) {
            uint32_t v;
            memcpy(&v, r + 4 * i, 4);
            out.push_back(v);
        }
    };
    put_root(t_main.tree->root);
    put_root(t_int.tree->root);
    put_root(qtree->root);
    uint32_t w4[4];
    qstore(cl, w4);
    out.insert(out.end(), w4, w4 + 4);
    qstore(ct, w4);
    out.insert(out.end(), w4, w4 + 4);
    out.insert(out.end(), sv.begin(), sv.end());
    out.push_back((uint32_t)gkr.size());
    out.insert(out.end(), gkr.begin(), gkr.end());
    put_root(ch.digest);
    c.sync();
    if (total_ms) *total_ms = lp_ms() - t0;
    return out;
}

}  // namespace sr
