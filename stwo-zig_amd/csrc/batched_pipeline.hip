// Batched primitive pipeline: many columns of uneven log sizes LDE'd and committed,
// then FRI fold_line rounds, planned once and replayed as one hipGraph.
//
// Inputs (seeded, generated at plan time): n_columns columns of uniform M31 values on
// CanonicCoset(log_i) (evaluations, bit-reversed storage), log_i drawn from
// [log_min, log_max]; one QM31 line evaluation of log fri_log.
// Replayed sequence (all launches, no host traffic):
//   1. interpolate every column (transforms.zig:249-360), out of place, per size class;
//   2. evaluate at blowup 2^log_blowup (transforms.zig:149-246);
//   3. one lifted Merkle commitment over all LDE columns (leaves.zig:24-200, layers.zig:99-140);
//   4. device channel <- main root; for each of n_folds rounds: commit the line (leaves,
//      parents), mix its root and draw alpha on the device, fold_line (folding.zig:352-380)
//      with the next layer's leaves fused; the last line is committed and mixed too.
// Result words: main root, n_folds + 1 line roots (8 words each), final channel digest (8)
// and draw counter.
//
// Tuning: fft_tile_log (LDS tile of the FFT passes) and merkle_levels (parent levels per
// launch) change only launch shapes; the autotuner times each candidate graph on the
// actual shapes, checks that every candidate produces the same result words, and keeps
// the fastest.  The shapes belong to the pipeline: the context's own settings are
// restored after every record.
#include <algorithm>
#include <vector>

#include "prover.hpp"

namespace sr {

__device__ __forceinline__ uint64_t bp_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// value(col, row) = splitmix64(seed + (col << 40) + row) mod p
__global__ void bp_gen_column_kernel(uint64_t seed, uint32_t col, uint64_t n, uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (uint32_t)(bp_splitmix64(seed + ((uint64_t)col << 40) + i) % P);
}
// line word k of element i = splitmix64((seed ^ 0x51F15EED) + 4i + k) mod p
__global__ void bp_gen_line_kernel(uint64_t seed, uint64_t n, uint32_t* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * n) return;
    out[i] = (uint32_t)(bp_splitmix64((seed ^ 0x51F15EEDull) + i) % P);
}

static uint64_t bp_splitmix64_host(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

uint32_t batched_column_log(const BatchedSpec& s, uint32_t i) {
    return s.log_min + (uint32_t)(bp_splitmix64_host(s.seed ^ 0xA5A5A5A5ull ^ (i * 0x9E3779B97F4A7C15ull)) %
                                  (s.log_max - s.log_min + 1));
}

BatchedPipeline::BatchedPipeline(Ctx& cc, const BatchedSpec& s) : c(cc), sp(s), owner(cc) {
    tile_log = c.fft_tile_log;
    levels = c.merkle_levels;
    SR_CHECK(sp.n_columns >= 1 && sp.n_columns <= 4096, "batched: n_columns");
    SR_CHECK(sp.log_min >= 1 && sp.log_min <= sp.log_max && sp.log_max + sp.log_blowup <= 28, "batched: log range");
    SR_CHECK(sp.log_blowup >= 1 && sp.log_blowup <= 4, "batched: log_blowup");
    SR_CHECK(sp.fri_log >= 2 && sp.fri_log <= 26 && sp.n_folds >= 1 && sp.n_folds < sp.fri_log, "batched: FRI shape");
    const uint32_t B = sp.log_blowup;
    max_lde = 0;
    logs.resize(sp.n_columns);
    uint64_t in_words = 0, lde_words = 0;
    for (uint32_t i = 0; i < sp.n_columns; i++) {
        logs[i] = batched_column_log(sp, i);
        in_words += 1ull << logs[i];
        lde_words += 1ull << (logs[i] + B);
        max_lde = std::max(max_lde, logs[i] + B);
    }
    c.ensure_tower(std::max(max_lde, sp.fri_log + 1));
    evals = DBuf(&c, 4 * in_words);
    coeffs = DBuf(&c, 4 * in_words);
    lde = DBuf(&c, 4 * lde_words);
    std::vector<uint32_t*> ev(sp.n_columns), cf(sp.n_columns), ld(sp.n_columns);
    {
        uint64_t a = 0, b = 0;
        for (uint32_t i = 0; i < sp.n_columns; i++) {
            ev[i] = evals.as<uint32_t>() + a;
            cf[i] = coeffs.as<uint32_t>() + a;
            ld[i] = lde.as<uint32_t>() + b;
            a += 1ull << logs[i];
            b += 1ull << (logs[i] + B);
            const uint64_t n = 1ull << logs[i];
            hipLaunchKernelGGL(bp_gen_column_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream, sp.seed,
                               i, n, ev[i]);
        }
    }
    // size classes: one FFT launch chain per distinct log
    for (uint32_t lg = sp.log_min; lg <= sp.log_max; lg++) {
        std::vector<uint32_t*> s1, d1, d2;
        for (uint32_t i = 0; i < sp.n_columns; i++)
            if (logs[i] == lg) {
                s1.push_back(ev[i]);
                d1.push_back(cf[i]);
                d2.push_back(ld[i]);
            }
        if (s1.empty()) continue;
        SizeClass sc;
        sc.log = lg;
        sc.n = (uint32_t)s1.size();
        sc.d_eval = upload(c, s1.data(), s1.size());
        sc.d_coeff = upload(c, d1.data(), d1.size());
        sc.d_lde = upload(c, d2.data(), d2.size());
        classes.push_back(std::move(sc));
    }
    // lifted Merkle over the LDE columns sorted by (log, index) (columns.zig:12-46)
    std::vector<uint32_t> order(sp.n_columns);
    for (uint32_t i = 0; i < sp.n_columns; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return logs[x] < logs[y]; });
    std::vector<const uint32_t*> sc(sp.n_columns);
    std::vector<uint32_t> sl(sp.n_columns);
    for (uint32_t i = 0; i < sp.n_columns; i++) {
        sc[i] = ld[order[i]];
        sl[i] = logs[order[i]] + B;
    }
    LeafPlan plan = merkle_leaf_plan(sc.data(), sl.data(), sp.n_columns);
    leaf_tab = upload(c, plan.tab.data(), plan.tab.size());
    leaf_segs = plan.segs;
    tree = owner.make(max_lde);
    // FRI lines and trees
    lines.reserve(sp.n_folds + 1);
    for (uint32_t j = 0; j <= sp.n_folds; j++) {
        lines.emplace_back(&c, 16ull << (sp.fri_log - j));
        ftrees.push_back(owner.make(sp.fri_log - j));
    }
    {
        const uint64_t n = 1ull << sp.fri_log;
        hipLaunchKernelGGL(bp_gen_line_kernel, dim3((uint32_t)((4 * n + 255) / 256)), dim3(256), 0, c.stream, sp.seed,
                           n, lines[0].as<uint32_t>());
        SR_HIP(hipGetLastError());
    }
    chan = DBuf(&c, CH_BYTES + 32 * (sp.n_folds + 2));
    c.sync();
}

BatchedPipeline::~BatchedPipeline() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
}

// Applies a pipeline's launch shapes to the context for the scope of one record.
struct TuningScope {
    Ctx& c;
    uint32_t tile, levels;
    TuningScope(Ctx& cc, uint32_t t, uint32_t l) : c(cc), tile(cc.fft_tile_log), levels(cc.merkle_levels) {
        c.fft_tile_log = t;
        c.merkle_levels = l;
    }
    ~TuningScope() {
        c.fft_tile_log = tile;
        c.merkle_levels = levels;
    }
};

void BatchedPipeline::record() {
    TuningScope ts(c, tile_log, levels);
    const uint32_t B = sp.log_blowup;
    for (auto& sc : classes)
        launch_circle_fft(c, true, sc.d_eval.as<const uint32_t* const>(), sc.d_coeff.as<uint32_t* const>(), sc.n,
                          sc.log, sc.log);
    for (auto& sc : classes)
        launch_circle_fft(c, false, sc.d_coeff.as<const uint32_t* const>(), sc.d_lde.as<uint32_t* const>(), sc.n,
                          sc.log + B, sc.log);
    launch_merkle_leaves_dev(c, leaf_tab.as<LeafColDev>(), leaf_segs, sp.n_columns, *tree);
    launch_merkle_parents(c, *tree);
    DevChannel* d_ch = chan.as<DevChannel>();
    uint8_t* d_roots = chan.as<uint8_t>() + CH_BYTES;
    launch_fill_u32(c, chan.as<uint32_t>(), CH_BYTES / 4, 0);
    launch_chan_mix_root_draw(c, d_ch, tree->layers, d_roots);
    launch_merkle_leaves_qm31(c, lines[0].as<uint32_t>(), sp.fri_log, ftrees[0]->layer(sp.fri_log));
    launch_merkle_parents(c, *ftrees[0], d_ch, d_roots + 32);
    for (uint32_t j = 1; j <= sp.n_folds; j++)
        launch_fri_commit_line(c, lines[j - 1].as<uint32_t>(), sp.fri_log - j + 1, lines[j].as<uint32_t>(), *ftrees[j],
                               d_ch, d_roots + 32 * (1 + j));
}

void BatchedPipeline::capture() {
    if (exec) SR_HIP(hipGraphExecDestroy(exec));
    if (graph) SR_HIP(hipGraphDestroy(graph));
    exec = nullptr;
    graph = nullptr;
    const bool prof = c.profiling;
    c.profiling = false;  // no event nodes in the graph
    SR_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    try {
        record();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c.stream, &g);
        if (g) (void)hipGraphDestroy(g);
        c.profiling = prof;
        throw;
    }
    SR_HIP(hipStreamEndCapture(c.stream, &graph));
    c.profiling = prof;
    SR_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    cap_tower[0] = c.d_tw;
    cap_tower[1] = c.d_itw;
    cap_tower[2] = c.d_tw2;
    cap_tower[3] = c.d_itw2;
    cap_tower_log = c.tower_log;
}

bool BatchedPipeline::tower_moved() const {
    return cap_tower[0] != c.d_tw || cap_tower[1] != c.d_itw || cap_tower[2] != c.d_tw2 || cap_tower[3] != c.d_itw2 ||
           cap_tower_log != c.tower_log;
}

void BatchedPipeline::set_tuning(uint32_t t, uint32_t l) {
    tile_log = t;
    levels = l;
    capture();
}

double BatchedPipeline::replay(uint32_t n) {
    SR_CHECK(exec, "batched: graph not captured");
    if (tower_moved()) capture();
    hipEvent_t a, b;
    SR_HIP(hipEventCreate(&a));
    SR_HIP(hipEventCreate(&b));
    SR_HIP(hipEventRecord(a, c.stream));
    for (uint32_t i = 0; i < n; i++) SR_HIP(hipGraphLaunch(exec, c.stream));
    SR_HIP(hipEventRecord(b, c.stream));
    SR_HIP(hipEventSynchronize(b));
    float ms = 0;
    SR_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return n ? ms / n : 0.0;
}

uint64_t BatchedPipeline::input_words() const {
    uint64_t w = 0;
    for (uint32_t lg : logs) w += 1ull << lg;
    return w;
}

std::vector<uint32_t> BatchedPipeline::result() {
    const size_t bytes = CH_BYTES + 32 * (sp.n_folds + 2);
    uint8_t* h = c.pinned(bytes);
    SR_HIP(hipMemcpyAsync(h, chan.p, bytes, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    DevChannel dc;
    memcpy(&dc, h, sizeof dc);
    SR_CHECK(dc.status == 0, "batched: channel draw rejected 64 times in a row");
    std::vector<uint32_t> out;
    const uint32_t* r = (const uint32_t*)(h + CH_BYTES);
    out.insert(out.end(), r, r + 8 * (sp.n_folds + 2));
    out.insert(out.end(), dc.digest, dc.digest + 8);
    out.push_back(dc.n_draws);
    return out;
}

void BatchedPipeline::autotune(uint32_t reps) {
    const uint32_t tiles[3] = {11, 12, 13}, levels_set[3] = {6, 8, 10};
    double best = 1e30;
    uint32_t bt = tile_log, bl = levels;
    std::vector<uint32_t> ref;
    tuning.clear();
    for (uint32_t t : tiles)
        for (uint32_t l : levels_set) {
            tile_log = t;
            levels = l;
            capture();
            replay(1);  // warm
            const double ms = replay(reps);
            std::vector<uint32_t> res = result();
            if (ref.empty()) ref = res;
            SR_CHECK(res == ref, "batched: tuning candidates disagree");
            tuning.push_back({t, l, ms});
            if (ms < best) {
                best = ms;
                bt = t;
                bl = l;
            }
        }
    set_tuning(bt, bl);
}

}  // namespace sr
