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
// The host only seeds the channel with (n, seed): every later transcript step (mixRoot,
// the draws, the claimed sums, the GKR proof, the OODS point, the sampled values and the
// quotient coefficient) runs on a device channel (transcript.hip, gkr.hip), and the host
// reads the roots, sums, sampled values, GKR proof and final channel once at the end.
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
    const uint64_t N = 1ull << n, LN = 1ull << L;
    const uint32_t NS = 12;  // sampled polynomials: a, b, c, m and the 8 interaction coordinates
    c.ensure_tower(L + 1);
    TreeOwner owner(c);
    Channel ch;
    {
        uint32_t st[3] = {n, (uint32_t)seed, (uint32_t)(seed >> 32)};
        ch.mix_u32s(st, 3);
    }
    // From here on the transcript runs on a device copy of the channel.  Block `blk` holds the
    // readback region [DevChannel | roots of the main, interaction and quotient trees |
    // claimed sums | zero-denominator flags | z, alpha | OODS point and evaluation factors |
    // sampled values], then the device tables of every launch; its host image goes up in one
    // copy before the first launch and the readback region comes back once at the end.
    const size_t at_roots = CH_BYTES, at_cs = at_roots + 96, at_flags = at_cs + 32, at_za = at_flags + 16;
    const size_t at_oods = at_za + 32, at_sv = at_oods + 16 * (2 + 32), rb_bytes = at_sv + 16 * (size_t)NS;
    HostPack pack;
    {
        DevChannel hc{};
        memcpy(hc.digest, ch.digest, 32);
        hc.n_draws = ch.n_draws;
        pack.add(&hc, 1);  // offset 0
        pack.buf.resize(rb_bytes, 0);
    }
    // 1. trace columns a, b, c, m (evaluations, kept for the denominators) and coefficients
    DBuf main(&c, 4 * N * 4);
    uint32_t* col[4];
    for (int k = 0; k < 4; k++) col[k] = main.as<uint32_t>() + k * N;
    DBuf mult16(&c, 4u << 16), negm(&c, 4 * N);
    DBuf main_cf(&c, 4 * N * 4);
    std::vector<const uint32_t*> mcf(4);
    for (int k = 0; k < 4; k++) mcf[k] = main_cf.as<uint32_t>() + k * N;
    const size_t at_ifft_src = pack.add(col, 4), at_ifft_dst = pack.add(mcf.data(), 4);
    Committed t_main;
    CommitPlan cp_main;
    plan_commit(c, owner, t_main, mcf, n, B, pack, cp_main);
    // 3-5. denominators, cumulative columns, interaction trace (8 coordinate columns)
    DBuf dl(&c, 16 * N), dt(&c, 16 * N), sl(&c, 16 * N), stt(&c, 16 * N);
    DBuf inter(&c, 8 * N * 4);
    uint32_t* ic[8];
    for (int k = 0; k < 8; k++) ic[k] = inter.as<uint32_t>() + k * N;
    std::vector<const uint32_t*> icf(ic, ic + 8);
    const size_t at_ic = pack.add(ic, 8);
    Committed t_int;
    CommitPlan cp_int;
    plan_commit(c, owner, t_int, icf, n, B, pack, cp_int);
    // 7-8. sampled columns and quotient entries (every LDE column has the quotient domain's size)
    std::vector<const uint32_t*> all_cf(mcf);
    all_cf.insert(all_cf.end(), icf.begin(), icf.end());
    std::vector<uint32_t*> all_lde(t_main.lde);
    all_lde.insert(all_lde.end(), t_int.lde.begin(), t_int.lde.end());
    SR_CHECK(all_cf.size() == NS && all_lde.size() == NS, "logup: sampled column count");
    const size_t at_cols = pack.add(all_cf.data(), NS);
    size_t at_ents;
    {
        std::vector<QEntryDev> ents(NS);
        for (uint32_t i = 0; i < NS; i++) {
            ents[i].col = all_lde[i];
            ents[i].shift = 1;
            ents[i].pad = 0;
            for (int k = 0; k < 4; k++) ents[i].coeff[k] = 0;  // written by the transcript step
        }
        at_ents = pack.add(ents.data(), NS);
    }
    DBuf quot(&c, 4 * LN * 4);
    uint32_t* qc[4];
    for (int k = 0; k < 4; k++) qc[k] = quot.as<uint32_t>() + k * LN;
    DeviceTree* qtree = owner.make(L);
    DBuf blk = upload(c, pack.buf.data(), pack.buf.size());
    uint8_t* d_pack = blk.as<uint8_t>();
    DevChannel* d_ch = (DevChannel*)d_pack;
    uint8_t* d_roots = d_pack + at_roots;
    uint32_t* d_cs = (uint32_t*)(d_pack + at_cs);
    uint32_t* d_flags = (uint32_t*)(d_pack + at_flags);
    uint32_t* d_za = (uint32_t*)(d_pack + at_za);
    uint32_t* d_oods = (uint32_t*)(d_pack + at_oods);
    uint32_t* d_factors = d_oods + 8;
    uint32_t* d_sv = (uint32_t*)(d_pack + at_sv);

    // 1-2. trace, commitment, mixRoot; z, alpha (commit_polys.zig:10-67)
    launch_xor_trace(c, n, seed, col[0], col[1], col[2], mult16.as<uint32_t>());
    launch_logup_table_cols(c, mult16.as<uint32_t>(), N, col[3], negm.as<uint32_t>());
    launch_circle_fft(c, true, (const uint32_t* const*)(d_pack + at_ifft_src), (uint32_t* const*)(d_pack + at_ifft_dst),
                      4, n, n);
    launch_commit(c, t_main, cp_main, d_pack, d_roots);
    launch_tr_mix_draws(c, d_ch, (const uint32_t*)d_roots, 8, 2, d_za);
    // 3. denominators
    launch_logup_denominators_dev(c, col[0], col[1], col[2], N, d_za, dl.as<uint32_t>());
    launch_logup_denominators_dev(c, nullptr, nullptr, nullptr, N, d_za, dt.as<uint32_t>());
    // 4. cumulative columns; the claimed sums are mixed (their sum is checked after the readback)
    logup_cumulative_dev(c, 2, nullptr, dl.as<uint32_t>(), N, sl.as<uint32_t>(), d_cs, d_flags);
    logup_cumulative_dev(c, 1, negm.as<uint32_t>(), dt.as<uint32_t>(), N, stt.as<uint32_t>(), d_cs + 4, d_flags + 1);
    launch_tr_mix_draws(c, d_ch, d_cs, 8, 0, nullptr);
    // 5. interaction commitment, mixRoot
    launch_qm31_to_coords(c, sl.as<uint32_t>(), N, ic);
    launch_qm31_to_coords(c, stt.as<uint32_t>(), N, ic + 4);
    launch_circle_fft(c, true, (const uint32_t* const*)(d_pack + at_ic), (uint32_t* const*)(d_pack + at_ic), 8, n, n);
    launch_commit(c, t_int, cp_int, d_pack, d_roots + 32);
    launch_tr_mix_draws(c, d_ch, (const uint32_t*)(d_roots + 32), 8, 0, nullptr);
    // 6. GKR batch proof of the two fraction sums on the device channel
    GkrRun* gkr_run = gkr_launch_on_channel(
        c, d_ch,
        {GkrInstance{GKR_LOGUP_SINGLES, nullptr, dl.as<uint32_t>(), n},
         GkrInstance{GKR_LOGUP_MULTIPLICITIES, negm.as<uint32_t>(), dt.as<uint32_t>(), n}});
    std::vector<uint32_t> gkr;
    try {
        // 7. OODS point (circle.zig:159-167) and the sampled values (one weight vector)
        launch_tr_oods(c, d_ch, nullptr, n, d_oods, d_factors);
        {
            DBuf w(&c, N * 16);
            launch_eval_weights_dev(c, d_factors, n, w.as<uint32_t>());
            launch_eval_columns(c, (const uint32_t* const*)(d_pack + at_cols), NS, n, w.as<uint32_t>(), d_sv);
        }
        // 8. mixFelts(sampled values), quotient coefficient and batch, quotients, commitment
        QEntryDev* d_ents = (QEntryDev*)(d_pack + at_ents);
        DBuf d_bat(&c, sizeof(QuotientBatchDev));
        launch_tr_quotient(c, d_ch, d_sv, NS, d_oods, d_ents, d_bat.as<QuotientBatchDev>());
        launch_quotients_dev(c, d_ents, NS, d_bat.as<QuotientBatchDev>(), 1, 4.0 * NS * (double)LN, L, qc,
                             qtree->layer(L), 1, 0, true);
        launch_merkle_parents(c, *qtree, nullptr, d_roots + 64);
        launch_tr_mix_draws(c, d_ch, (const uint32_t*)(d_roots + 64), 8, 0, nullptr);
    } catch (...) {
        gkr_release(gkr_run);
        throw;
    }
    gkr = gkr_collect(c, gkr_run);  // one synchronisation for the whole workload
    uint8_t* h = c.pinned(rb_bytes);
    SR_HIP(hipMemcpyAsync(h, d_pack, rb_bytes, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    DevChannel hc;
    memcpy(&hc, h, sizeof(hc));
    SR_CHECK(hc.status == 0, "logup: channel draw rejected 64 times in a row");
    const uint32_t* flags = (const uint32_t*)(h + at_flags);
    SR_CHECK(flags[0] == 0 && flags[1] == 0, "LogUp: zero denominator");
    const qm31 cl = qload((const uint32_t*)(h + at_cs)), ct = qload((const uint32_t*)(h + at_cs) + 4);
    SR_CHECK(qiszero(qadd(cl, ct)), "logup: lookups and table do not balance");

    std::vector<uint32_t> out;
    auto put_words = [&](const uint8_t* p, size_t n_words) {
        for (size_t i = 0; i < n_words; i++) {
            uint32_t v;
            memcpy(&v, p + 4 * i, 4);
            out.push_back(v);
        }
    };
    put_words(h + at_roots, 24);  // main, interaction and quotient roots
    put_words(h + at_cs, 8);      // claimed sums
    put_words(h + at_sv, 4 * NS);
    out.push_back((uint32_t)gkr.size());
    out.insert(out.end(), gkr.begin(), gkr.end());
    put_words(h, 8);  // final channel digest
    if (total_ms) *total_ms = lp_ms() - t0;
    return out;
}

}  // namespace sr
