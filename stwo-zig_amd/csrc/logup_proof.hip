// The XOR-lookup AIR as a whole STARK proof on one GPU.
//
// The AIR is defined by this backend (its LogUp benchmark), with the LogUp machinery of the
// reference's RISC-V frontend: the pair-batched cumulative column and transition constraint of
// src/frontends/riscv/air/logup.zig:74-105 over a lookup-table component shaped like
// src/frontends/riscv/air/lookups/tables/component.zig (is_first and the table tuple
// preprocessed, the multiplicities committed, one interaction column sampled at the OODS point
// and at the previous row).
//   tree 0 (preprocessed): is_first, tx, ty, tz = ((i >> 8) & 255, i & 255, tx ^ ty)
//   tree 1 (main): a, b uniform 8-bit, c = a ^ b, m = table multiplicities (zero past 2^16)
//   tree 2 (interaction): S_i = sum_{j <= i} (1 / d1_j - m_j / d2_j) in coset row order,
//          d1 = z - (a + alpha b + alpha^2 c), d2 = z - (tx + alpha ty + alpha^2 tz)
//   tree 3: composition of (S(x) - S(x - g) + is_first(x) claimed) d1 d2 - (d2 - m d1)
//          over the trace coset's vanishing polynomial, g the coset step (x - g: the previous row)
// Row i of a column is stored at the bit-reversed circle-domain position of coset index i, as
// for the reference's periodic XOR example.  The lookups balance the table, so the claimed sum
// is zero.  Transcript (prover_transaction.zig:222-339 order): config, root 0, root 1,
// mixU32s([log_n]), z, alpha, root 2, mixFelts([claimed]), composition coefficient, root 3, OODS
// point, sampled values, then the rest of the proof as for every AIR (finish_proof).
#include <algorithm>
#include <vector>

#include "prover.hpp"

namespace sr {

namespace {

__device__ __forceinline__ uint32_t lp_position(uint32_t log_size, uint32_t i) {
    const uint32_t n = 1u << log_size;
    const uint32_t ci = (i & 1) == 0 ? i / 2 : (2 * n - i) / 2;
    return bitrev32(ci, log_size);
}

__device__ __forceinline__ qm31 lp_combine(qm31 z, qm31 al, qm31 al2, uint32_t x, uint32_t y, uint32_t w) {
    return qsub(z, qadd(qadd(qfromm(x), qmulm(al, y)), qmulm(al2, w)));
}

// Committed columns at their positions: pre = is_first, tx, ty, tz; main = a, b, c, m.
__global__ __launch_bounds__(256) void lp_columns_kernel(uint32_t log_size, const uint32_t* __restrict__ ar,
                                                         const uint32_t* __restrict__ br,
                                                         const uint32_t* __restrict__ cr,
                                                         const uint32_t* __restrict__ mult16,
                                                         uint32_t* __restrict__ pre, uint32_t* __restrict__ mainv) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (i >= n) return;
    const uint32_t p = lp_position(log_size, i);
    const uint32_t tx = (i >> 8) & 255u, ty = i & 255u;
    pre[p] = i == 0 ? 1u : 0u;
    pre[n + p] = tx;
    pre[2 * n + p] = ty;
    pre[3 * n + p] = tx ^ ty;
    mainv[p] = ar[i];
    mainv[n + p] = br[i];
    mainv[2 * n + p] = cr[i];
    mainv[3 * n + p] = i < 65536u ? mult16[i] : 0u;
}

// Row pair fraction (1 / d1) + (-m / d2) as num / den = (d2 - m d1) / (d1 d2), rows in coset order.
__global__ __launch_bounds__(256) void lp_pair_kernel(uint32_t n, const uint32_t* __restrict__ ar,
                                                      const uint32_t* __restrict__ br, const uint32_t* __restrict__ cr,
                                                      const uint32_t* __restrict__ mult16, qm31 z, qm31 al, qm31 al2,
                                                      uint4* __restrict__ num, uint4* __restrict__ den) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t tx = (i >> 8) & 255u, ty = i & 255u;
    const qm31 d1 = lp_combine(z, al, al2, ar[i], br[i], cr[i]);
    const qm31 d2 = lp_combine(z, al, al2, tx, ty, tx ^ ty);
    const uint32_t m = i < 65536u ? mult16[i] : 0u;
    const qm31 nu = qsub(d2, qmulm(d1, m)), de = qmul(d1, d2);
    num[i] = make_uint4(nu.c0.a, nu.c0.b, nu.c1.a, nu.c1.b);
    den[i] = make_uint4(de.c0.a, de.c0.b, de.c1.a, de.c1.b);
}

// Cumulative sums (coset order, QM31 AoS) to the 4 coordinate columns at their positions.
__global__ __launch_bounds__(256) void lp_scatter_kernel(uint32_t log_size, const uint4* __restrict__ sums,
                                                         uint32_t* __restrict__ coords) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (i >= n) return;
    const uint32_t p = lp_position(log_size, i);
    const uint4 v = sums[i];
    coords[p] = v.x;
    coords[n + p] = v.y;
    coords[2 * n + p] = v.z;
    coords[3 * n + p] = v.w;
}

// Composition on CanonicCoset(n + 1) (bit-reversed positions r): ev[0..12) = is_first, tx, ty,
// tz, a, b, c, m, S0..S3.  S(x - g) is read at the position of x - g: the natural index i of r
// moves one step back on the half coset (i - 1) and one forward on its conjugate.
__global__ __launch_bounds__(256) void lp_composition_kernel(uint32_t log_n, const uint32_t* const* __restrict__ ev,
                                                             qm31 z, qm31 al, qm31 al2, qm31 claimed, uint32_t den0,
                                                             uint32_t den1, uint32_t* __restrict__ o0,
                                                             uint32_t* __restrict__ o1, uint32_t* __restrict__ o2,
                                                             uint32_t* __restrict__ o3) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t lg = log_n + 1, half = 1u << log_n;
    if (r >= (2u << log_n)) return;
    const uint32_t i = bitrev32(r, lg);
    const uint32_t j = i < half ? ((i + half - 1) & (half - 1)) : (half + ((i - half + 1) & (half - 1)));
    const uint32_t q = bitrev32(j, lg);
    const qm31 S = {{ev[8][r], ev[9][r]}, {ev[10][r], ev[11][r]}};
    const qm31 Sp = {{ev[8][q], ev[9][q]}, {ev[10][q], ev[11][q]}};
    const qm31 d1 = lp_combine(z, al, al2, ev[4][r], ev[5][r], ev[6][r]);
    const qm31 d2 = lp_combine(z, al, al2, ev[1][r], ev[2][r], ev[3][r]);
    const qm31 delta = qadd(qsub(S, Sp), qmulm(claimed, ev[0][r]));
    qm31 cons = qsub(qmul(qmul(delta, d1), d2), qsub(d2, qmulm(d1, ev[7][r])));
    cons = qmulm(cons, (r >> log_n) ? den1 : den0);
    o0[r] = cons.c0.a;
    o1[r] = cons.c0.b;
    o2[r] = cons.c1.a;
    o3[r] = cons.c1.b;
}

// Evaluations (device columns of 2^n) -> coefficients (new buffer) -> LDE -> tree; root read back.
void commit_evals(Ctx& c, TreeOwner& owner, Committed& out, DBuf& coeff_store, const std::vector<const uint32_t*>& ev,
                  uint32_t n, uint32_t B) {
    const uint64_t N = 1ull << n;
    const uint32_t k = (uint32_t)ev.size();
    coeff_store = DBuf(&c, 4 * N * k);
    std::vector<const uint32_t*> cf(k);
    for (uint32_t i = 0; i < k; i++) cf[i] = coeff_store.as<uint32_t>() + i * N;
    HostPack pack;
    CommitPlan cp;
    plan_commit(c, owner, out, cf, n, B, pack, cp);
    const size_t at_ev = pack.add(ev.data(), k);
    DBuf blk = upload(c, pack.buf.data(), pack.buf.size());
    launch_commit_from_evals(c, out, cp, blk.as<uint8_t>(), at_ev, nullptr);
    tree_read_root(c, *out.tree);
}

uint32_t vanishing_inv(Ctx& c, uint32_t n, uint32_t r) {
    // cosetVanishing(CanonicCoset(n).coset(), CanonicCoset(n + 1).circleDomain().at(r))
    // (core/constraints.zig:16-27)
    pt dp = idx_to_point(c.gen, domain_index_at(n + 1, (uint64_t)r));
    pt init = idx_to_point(c.gen, subgroup_gen(n + 1));
    pt s = padd(padd(dp, pconj(init)), init);
    uint32_t x = s.x;
    for (uint32_t i = 1; i < n; i++) x = mdouble_x(x);
    return minv(x);
}

}  // namespace

void prove_logup_xor(Ctx& c, uint32_t log_n, uint64_t seed, const ProveParams& prm, std::string& proof) {
    const uint32_t n = log_n, B = prm.log_blowup_factor;
    SR_CHECK(n >= 16 && n <= 26, "XOR-lookup AIR: log_n must be in 16..26");
    SR_CHECK(B >= 1 && B <= 4, "log_blowup_factor out of range");
    SR_CHECK(prm.n_queries >= 1 && prm.n_queries <= 1024, "n_queries out of range");
    SR_CHECK(prm.log_last_layer_degree_bound <= 10, "log_last_layer_degree_bound out of range");
    SR_CHECK(prm.pow_bits <= 40, "pow_bits out of range");
    const uint32_t L = n + B, last_log = prm.log_last_layer_degree_bound + B;
    SR_CHECK(L > last_log, "trace too small for the FRI configuration");
    const uint64_t N = 1ull << n, LN = 1ull << L;
    c.ensure_tower(std::max(L, n + 1) + 1);
    TreeOwner owner(c);
    Channel ch;
    qm31 cfg = {{prm.pow_bits, B}, {prm.n_queries, prm.log_last_layer_degree_bound}};
    ch.mix_felts(&cfg, 1);

    // trace rows (coset order), then the committed columns at their positions
    DBuf rows(&c, 12 * N), mult16(&c, 4u << 16);
    uint32_t* ar = rows.as<uint32_t>();
    uint32_t *br = ar + N, *cr = ar + 2 * N;
    launch_xor_trace(c, n, seed, ar, br, cr, mult16.as<uint32_t>());
    DBuf pre(&c, 16 * N), mainv(&c, 16 * N);
    hipLaunchKernelGGL(lp_columns_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, c.stream, n, ar, br, cr,
                       mult16.as<uint32_t>(), pre.as<uint32_t>(), mainv.as<uint32_t>());
    SR_HIP(hipGetLastError());
    auto cols4 = [&](const DBuf& b) {
        std::vector<const uint32_t*> v(4);
        for (int k = 0; k < 4; k++) v[k] = b.as<uint32_t>() + k * N;
        return v;
    };
    Committed t0, t1, t2, t3;
    DBuf cf0, cf1, cf2;
    commit_evals(c, owner, t0, cf0, cols4(pre), n, B);
    commit_evals(c, owner, t1, cf1, cols4(mainv), n, B);
    pre.reset();
    mainv.reset();
    ch.mix_root(t0.tree->root);
    ch.mix_root(t1.tree->root);
    ch.mix_u32s(&n, 1);
    const qm31 z = ch.draw_secure_felt(), al = ch.draw_secure_felt(), al2 = qsqr(al);

    // interaction column: pair fractions, prefix sums, coordinates at their positions
    qm31 claimed;
    {
        DBuf num(&c, 16 * N), den(&c, 16 * N), sums(&c, 16 * N), inter(&c, 16 * N);
        hipLaunchKernelGGL(lp_pair_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, c.stream, (uint32_t)N,
                           ar, br, cr, mult16.as<uint32_t>(), z, al, al2, num.as<uint4>(), den.as<uint4>());
        SR_HIP(hipGetLastError());
        claimed = logup_cumulative(c, 0, num.as<uint32_t>(), den.as<uint32_t>(), N, sums.as<uint32_t>());
        SR_CHECK(qiszero(claimed), "XOR-lookup AIR: the lookups do not balance the table");
        hipLaunchKernelGGL(lp_scatter_kernel, dim3((uint32_t)((N + 255) / 256)), dim3(256), 0, c.stream, n,
                           sums.as<uint4>(), inter.as<uint32_t>());
        SR_HIP(hipGetLastError());
        commit_evals(c, owner, t2, cf2, cols4(inter), n, B);
    }
    rows.reset();
    mult16.reset();
    ch.mix_root(t2.tree->root);
    ch.mix_felts(&claimed, 1);
    (void)ch.draw_secure_felt();  // composition coefficient: one constraint, power 1

    // composition on CanonicCoset(n + 1): the LDEs themselves at blowup 2, else own evaluations
    const uint64_t CN = 2 * N;
    DBuf comp(&c, 16 * CN);
    uint32_t* cc[4];
    for (int k = 0; k < 4; k++) cc[k] = comp.as<uint32_t>() + k * CN;
    {
        const Committed* tr3[3] = {&t0, &t1, &t2};
        std::vector<const uint32_t*> ev(12);
        DBuf own;
        if (B == 1) {
            for (int k = 0; k < 12; k++) ev[k] = tr3[k / 4]->lde[k % 4];
        } else {
            own = DBuf(&c, 4 * CN * 12);
            std::vector<const uint32_t*> src(12);
            std::vector<uint32_t*> dst(12);
            for (int k = 0; k < 12; k++) {
                src[k] = tr3[k / 4]->coeffs[k % 4];
                dst[k] = own.as<uint32_t>() + k * CN;
                ev[k] = dst[k];
            }
            DBuf ds = upload(c, src.data(), 12), dd = upload(c, dst.data(), 12);
            launch_circle_fft(c, false, ds.as<const uint32_t* const>(), dd.as<uint32_t* const>(), 12, n + 1, n);
            c.sync();
        }
        DBuf dev = upload(c, ev.data(), 12);
        KTimer kt(c, "logup_composition", 4.0 * 13 * (double)CN + 16.0 * (double)CN);
        hipLaunchKernelGGL(lp_composition_kernel, dim3((uint32_t)((CN + 255) / 256)), dim3(256), 0, c.stream, n,
                           dev.as<const uint32_t* const>(), z, al, al2, claimed, vanishing_inv(c, n, 0),
                           vanishing_inv(c, n, 1), cc[0], cc[1], cc[2], cc[3]);
        SR_HIP(hipGetLastError());
        DBuf dcc = upload(c, cc, 4);
        launch_circle_fft(c, true, dcc.as<const uint32_t* const>(), dcc.as<uint32_t* const>(), 4, n + 1, n + 1);
        c.sync();
    }
    {
        std::vector<const uint32_t*> halves(8);
        for (int k = 0; k < 4; k++) {
            halves[k] = cc[k];
            halves[4 + k] = cc[k] + N;
        }
        commit_coefficients(c, owner, t3, halves, n, B);
        tree_read_root(c, *t3.tree);
    }
    ch.mix_root(t3.tree->root);

    // OODS point (circle.zig:159-167), the previous-row point, sampled values
    const qm31 tt = ch.draw_secure_felt(), t2q = qsqr(tt);
    const qm31 inv1 = qinv(qadd(t2q, qone()));
    const SamplePoint oods = {qmul(qsub(qone(), t2q), inv1), qmul(qadd(tt, tt), inv1)};
    SamplePoint prev;
    {
        const pt g = idx_to_point(c.gen, (uint32_t)((0x80000000u - subgroup_gen(n)) & ORDER_MASK));
        prev = {qsub(qmul(oods.x, qfromm(g.x)), qmul(oods.y, qfromm(g.y))),
                qadd(qmul(oods.x, qfromm(g.y)), qmul(oods.y, qfromm(g.x)))};
    }
    const Committed* trees[4] = {&t0, &t1, &t2, &t3};
    std::vector<const uint32_t*> all_cf;
    for (auto* T : trees) all_cf.insert(all_cf.end(), T->coeffs.begin(), T->coeffs.end());
    const uint32_t NC = (uint32_t)all_cf.size();  // 20
    auto eval_at = [&](const SamplePoint& p, const std::vector<const uint32_t*>& cols) {
        qm31 f[32];
        f[0] = p.y;
        f[1] = p.x;
        for (uint32_t k = 2; k < n; k++) {
            const qm31 sq = qsqr(f[k - 1]);
            f[k] = qsub(qadd(sq, sq), qone());
        }
        DBuf w(&c, N * 16), o(&c, 16 * cols.size());
        launch_eval_weights(c, f, n, w.as<uint32_t>());
        DBuf dcol = upload(c, cols.data(), cols.size());
        launch_eval_columns(c, dcol.as<const uint32_t* const>(), (uint32_t)cols.size(), n, w.as<uint32_t>(),
                            o.as<uint32_t>());
        std::vector<uint32_t> h(4 * cols.size());
        SR_HIP(hipMemcpyAsync(h.data(), o.p, 16 * cols.size(), hipMemcpyDeviceToHost, c.stream));
        c.sync();
        return h;
    };
    const std::vector<uint32_t> cur = eval_at(oods, all_cf);
    const std::vector<uint32_t> prv = eval_at(prev, t2.coeffs);
    // tree/column/sample order: the interaction columns carry (current, previous)
    std::vector<uint32_t> sv;
    std::vector<uint32_t> n_samples;
    std::vector<std::vector<ColumnSample>> samples;
    std::vector<uint32_t> col_logs;
    std::vector<const uint32_t*> qcols;
    for (uint32_t i = 0; i < NC; i++) {
        const bool inter = i >= 8 && i < 12;
        sv.insert(sv.end(), &cur[4 * i], &cur[4 * i] + 4);
        std::vector<ColumnSample> cs = {{oods, qload(&cur[4 * i])}};
        if (inter) {
            sv.insert(sv.end(), &prv[4 * (i - 8)], &prv[4 * (i - 8)] + 4);
            cs.push_back({prev, qload(&prv[4 * (i - 8)])});
        }
        n_samples.push_back(inter ? 2u : 1u);
        samples.push_back(cs);
        col_logs.push_back(L);
    }
    for (auto* T : trees) qcols.insert(qcols.end(), T->lde.begin(), T->lde.end());
    {
        std::vector<qm31> f(sv.size() / 4);
        for (size_t k = 0; k < f.size(); k++) f[k] = qload(&sv[4 * k]);
        ch.mix_felts(f.data(), f.size());
    }

    // quotients with the first layer's leaves, then FRI on a device copy of the channel
    const qm31 qa = ch.draw_secure_felt();
    QuotientPlan qp = plan_quotients(c, L, col_logs, samples, qa);
    DBuf quot(&c, 16 * LN);
    uint32_t* qc[4];
    for (int k = 0; k < 4; k++) qc[k] = quot.as<uint32_t>() + k * LN;
    DeviceTree* ftree0 = owner.make(L);
    launch_quotients(c, qcols.data(), col_logs.data(), qp.entry_cols.data(), qp.entry_coeffs.data(),
                     (uint32_t)qp.entry_cols.size(), qp.batches.data(), (uint32_t)qp.batches.size(), L, qc,
                     ftree0->layer(L));
    const uint32_t n_inner = L - 1 - last_log;
    const size_t at_line = CH_BYTES + 32 * (1 + (size_t)n_inner), rb = at_line + (16ull << last_log);
    DBuf blk(&c, rb);
    {
        std::vector<uint8_t> h(rb, 0);
        DevChannel hc{};
        memcpy(hc.digest, ch.digest, 32);
        hc.n_draws = ch.n_draws;
        memcpy(h.data(), &hc, sizeof hc);
        SR_HIP(hipMemcpyAsync(blk.p, c.stage(h.data(), rb), rb, hipMemcpyHostToDevice, c.stream));
    }
    uint8_t* d_blk = blk.as<uint8_t>();
    DevChannel* d_ch = (DevChannel*)d_blk;
    uint8_t* d_roots = d_blk + CH_BYTES;
    uint32_t* d_last = (uint32_t*)(d_blk + at_line);
    launch_merkle_parents(c, *ftree0, d_ch, d_roots);
    std::vector<DBuf> lines;
    std::vector<DeviceTree*> ftrees;
    const uint32_t* cq[4] = {qc[0], qc[1], qc[2], qc[3]};
    if (L - 1 == last_log) {
        launch_fold_circle_into_line(c, cq, L, qzero(), d_last, false, nullptr, d_ch);
    } else {
        for (uint32_t k = 0; k < n_inner; k++) lines.emplace_back(&c, (1ull << (L - 1 - k)) * 16);
        DeviceTree* t = owner.make(L - 1);
        launch_fri_commit_circle(c, cq, L, lines[0].as<uint32_t>(), *t, d_ch, d_roots + 32);
        ftrees.push_back(t);
        std::vector<uint32_t*> below;
        for (uint32_t k = 1; k < n_inner; k++) below.push_back(lines[k].as<uint32_t>());
        fri_fold_commit_down(c, owner, lines[0].as<uint32_t>(), L - 1, last_log, d_ch, below.data(), d_last,
                             d_roots + 64, ftrees);
    }
    std::vector<qm31> lv(1ull << last_log);
    {
        uint8_t* h = c.pinned(rb);
        SR_HIP(hipMemcpyAsync(h, d_blk, rb, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        DevChannel hc;
        memcpy(&hc, h, sizeof hc);
        SR_CHECK(hc.status == 0, "FRI: channel draw rejected 64 times in a row");
        memcpy(ch.digest, hc.digest, 32);
        ch.n_draws = hc.n_draws;
        memcpy(ftree0->root, h + CH_BYTES, 32);
        for (size_t k = 0; k < ftrees.size(); k++) memcpy(ftrees[k]->root, h + CH_BYTES + 32 * (1 + k), 32);
        const uint32_t* w = (const uint32_t*)(h + at_line);
        for (size_t i = 0; i < lv.size(); i++) lv[i] = qload(w + 4 * i);
    }
    ProofTail tail;
    tail.L = L;
    tail.last_log = last_log;
    tail.trees = {&t0, &t1, &t2, &t3};
    tail.n_samples = n_samples;
    tail.sv = std::move(sv);
    tail.ftree0 = ftree0;
    for (int k = 0; k < 4; k++) tail.qc[k] = qc[k];
    tail.ftrees = ftrees;
    for (auto& ln : lines) tail.lines.push_back(ln.as<uint32_t>());
    tail.last_values = std::move(lv);
    finish_proof(c, ch, prm, tail, proof, [] {});
}

}  // namespace sr
