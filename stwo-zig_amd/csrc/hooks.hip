// Resident hook surface of the C-ABI: the optional backend hooks the reference's generic
// prover probes with @hasDecl (src/prover/fri.zig:341-586, prover/pcs/commit_dispatch.zig:34-90,
// the MetalCommitBackend set in src/backends/metal/commit_backend.zig:120-800), each one call
// that leaves its results in HBM, plus a device-resident Blake2s channel for the steps that
// run between them.  Declared in include/stwo_rocm.h.
#include <algorithm>

#include "capi.hpp"

struct stwo_rocm_channel {
    DevChannel* d = nullptr;  // in the owning context's arena
};

namespace {

// Interpolates `cols` in place when `interpolate`, extends every column to log + log_blowup
// into ext_out and commits the extended columns as one lifted tree.
void prepare_and_commit(Ctx& c, uint32_t* const* cols, const uint32_t* log_sizes, uint32_t n_cols,
                        uint32_t log_blowup, uint32_t* const* ext_out, bool interpolate, stwo_rocm_tree** tree_out,
                        uint8_t* root_out, double* gpu_ms) {
    SR_CHECK(tree_out, "null output");
    SR_CHECK((cols && log_sizes && ext_out) || !n_cols, "null columns");
    SR_CHECK(log_blowup <= 8, "log blowup out of range");
    check_aligned((const void* const*)cols, n_cols, "prepare_and_commit");
    check_aligned((const void* const*)ext_out, n_cols, "prepare_and_commit");
    uint32_t max_ext = 0;
    std::vector<uint32_t> ext_logs(n_cols);
    for (uint32_t i = 0; i < n_cols; i++) {
        SR_CHECK(log_sizes[i] >= 1 && log_sizes[i] + log_blowup <= 30, "column log size out of range");
        ext_logs[i] = log_sizes[i] + log_blowup;
        max_ext = std::max(max_ext, ext_logs[i]);
    }
    if (max_ext >= 3) c.ensure_tower(max_ext);
    // one transform launch per distinct log size (the columns of a tree usually share one)
    std::vector<uint32_t> logs(log_sizes, log_sizes + n_cols);
    std::sort(logs.begin(), logs.end());
    logs.erase(std::unique(logs.begin(), logs.end()), logs.end());
    std::vector<DBuf> tables;
    DeviceTree* t = nullptr;
    double ms = timed(c, [&] {
        for (uint32_t lg : logs) {
            std::vector<uint32_t*> src, dst;
            for (uint32_t i = 0; i < n_cols; i++)
                if (log_sizes[i] == lg) {
                    src.push_back(cols[i]);
                    dst.push_back(ext_out[i]);
                }
            tables.push_back(upload(c, src.data(), src.size()));
            tables.push_back(upload(c, dst.data(), dst.size()));
            const uint32_t* const* ds = tables[tables.size() - 2].as<const uint32_t* const>();
            if (interpolate && log_blowup == 1) {
                // interpolation in place and the extension, fused where the pass plans allow
                launch_circle_ifft_lde(c, ds, tables[tables.size() - 2].as<uint32_t* const>(),
                                       tables.back().as<uint32_t* const>(), (uint32_t)src.size(), lg);
                continue;
            }
            if (interpolate)
                launch_circle_fft(c, true, ds, tables[tables.size() - 2].as<uint32_t* const>(), (uint32_t)src.size(),
                                  lg, lg);
            launch_circle_fft(c, false, ds, tables.back().as<uint32_t* const>(), (uint32_t)src.size(),
                              lg + log_blowup, lg);
        }
        std::vector<const uint32_t*> ev(ext_out, ext_out + n_cols);
        t = merkle_commit(c, ev, ext_logs);
    });
    if (n_cols) tree_read_root(c, *t);
    if (root_out) memcpy(root_out, t->root, 32);
    *tree_out = new stwo_rocm_tree{t};
    if (gpu_ms) *gpu_ms = ms;
}

DevChannel* chan_of(stwo_rocm_channel* ch) {
    SR_CHECK(ch && ch->d, "null channel");
    return ch->d;
}

void chan_mix_words(Ctx& c, stwo_rocm_channel* ch, const uint32_t* words, uint64_t n) {
    DevChannel* d = chan_of(ch);
    SR_CHECK(words || !n, "null words");
    SR_CHECK(n <= (1ull << 26), "channel: too many words in one mix");
    DBuf w;
    if (n) w = upload(c, words, (size_t)n);
    // an empty mix still rehashes the digest (digest <- Blake2s(digest))
    launch_tr_mix_draws(c, d, w.as<uint32_t>(), (uint32_t)n, 0, nullptr, true);
    c.sync();
}

void chan_check(Ctx& c, stwo_rocm_channel* ch) {
    DevChannel h;
    SR_HIP(hipMemcpyAsync(&h, ch->d, sizeof h, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    SR_CHECK(h.status == 0, "channel: a draw was rejected 64 times in a row");
}

}  // namespace

extern "C" {

int stwo_rocm_prepare_and_commit_owned(stwo_rocm_ctx* ctx, uint32_t* const* cols, const uint32_t* log_sizes,
                                       uint32_t n_cols, uint32_t log_blowup, uint32_t* const* ext_out,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out, double* gpu_ms) {
    return guarded(ctx, [&](Ctx& c) {
        prepare_and_commit(c, cols, log_sizes, n_cols, log_blowup, ext_out, true, tree_out, root_out, gpu_ms);
    });
}

int stwo_rocm_prepare_and_commit_polys(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, const uint32_t* log_sizes,
                                       uint32_t n_cols, uint32_t log_blowup, uint32_t* const* ext_out,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out, double* gpu_ms) {
    return guarded(ctx, [&](Ctx& c) {
        prepare_and_commit(c, (uint32_t* const*)coeffs, log_sizes, n_cols, log_blowup, ext_out, false, tree_out,
                           root_out, gpu_ms);
    });
}

int stwo_rocm_secure_column_from_line(stwo_rocm_ctx* ctx, const uint32_t* d_line, uint64_t n,
                                      uint32_t* const* coords) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_line && coords, "null argument");
        check_aligned((const void* const*)&d_line, 1, "secure_column_from_line");
        launch_qm31_to_coords(c, d_line, n, coords);
        c.sync();
    });
}

int stwo_rocm_commit_line(stwo_rocm_ctx* ctx, const uint32_t* d_line, uint32_t log_size, stwo_rocm_tree** tree_out,
                          uint8_t* root_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_line && tree_out, "null argument");
        SR_CHECK(log_size >= 1 && log_size <= 29, "log size out of range");
        check_aligned((const void* const*)&d_line, 1, "commit_line");
        DeviceTree* t = merkle_commit_qm31(c, d_line, log_size);
        tree_read_root(c, *t);
        if (root_out) memcpy(root_out, t->root, 32);
        *tree_out = new stwo_rocm_tree{t};
    });
}

int stwo_rocm_fri_fold_line_and_commit(stwo_rocm_ctx* ctx, const uint32_t* d_src, uint32_t log_src,
                                       const uint32_t alpha[4], uint32_t n_folds, uint32_t* d_dst,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_src && d_dst && alpha && tree_out, "null argument");
        SR_CHECK(n_folds >= 1 && log_src <= 29 && log_src >= n_folds + 1, "fold count out of range");
        check_aligned((const void* const*)&d_src, 1, "fri_fold_line_and_commit");
        check_aligned((const void* const*)&d_dst, 1, "fri_fold_line_and_commit");
        const uint32_t dst_log = log_src - n_folds;
        c.ensure_tower(log_src + 1);
        qm31 a = qin(alpha);
        TreeOwner owner(c);
        DeviceTree* t = owner.make(dst_log);
        std::vector<DBuf> mid;
        const uint32_t* src = d_src;
        for (uint32_t f = 0; f + 1 < n_folds; f++) {
            mid.emplace_back(&c, (1ull << (log_src - f - 1)) * 16);
            launch_fold_line(c, src, log_src - f, a, mid.back().as<uint32_t>(), nullptr);
            src = mid.back().as<uint32_t>();
            a = qsqr(a);
        }
        // the last fold hashes its folded values as the tree's leaves
        launch_fold_line(c, src, dst_log + 1, a, d_dst, t->layer(dst_log));
        launch_merkle_parents(c, *t);
        tree_read_root(c, *t);
        owner.trees.clear();  // the tree now belongs to the caller's handle
        if (root_out) memcpy(root_out, t->root, 32);
        *tree_out = new stwo_rocm_tree{t};
    });
}

int stwo_rocm_fri_commit_layers(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* d_line,
                                uint32_t log_size, uint32_t log_last_size, uint32_t* const* d_lines,
                                stwo_rocm_tree** trees_out, uint32_t* d_last_line) {
    return guarded(ctx, [&](Ctx& c) {
        DevChannel* dch = chan_of(ch);
        SR_CHECK(d_line && trees_out && d_last_line, "null argument");
        SR_CHECK(log_last_size >= 1 && log_size > log_last_size && log_size <= 29, "layer sizes out of range");
        const uint32_t n_layers = log_size - log_last_size;
        SR_CHECK(n_layers == 1 || d_lines, "null line buffers");
        check_aligned((const void* const*)&d_line, 1, "fri_commit_layers");
        check_aligned((const void* const*)&d_last_line, 1, "fri_commit_layers");
        if (n_layers > 1) check_aligned((const void* const*)d_lines, n_layers - 1, "fri_commit_layers");
        c.ensure_tower(log_size + 1);
        TreeOwner owner(c);
        std::vector<DeviceTree*> trees;
        DBuf d_roots(&c, 32ull * n_layers);
        // the top line: leaves, tree, then mixRoot and the first fold alpha on the device channel
        DeviceTree* top = owner.make(log_size);
        launch_merkle_leaves_qm31(c, d_line, log_size, top->layer(log_size));
        launch_merkle_parents(c, *top, dch, d_roots.as<uint8_t>());
        trees.push_back(top);
        fri_fold_commit_down(c, owner, d_line, log_size, log_last_size, dch, d_lines, d_last_line,
                             d_roots.as<uint8_t>() + 32, trees);
        SR_CHECK(trees.size() == n_layers, "FRI cascade: layer count");
        std::vector<uint8_t> roots(32ull * n_layers);
        SR_HIP(hipMemcpyAsync(roots.data(), d_roots.p, roots.size(), hipMemcpyDeviceToHost, c.stream));
        chan_check(c, ch);
        for (uint32_t k = 0; k < n_layers; k++) {
            memcpy(trees[k]->root, roots.data() + 32ull * k, 32);
            trees_out[k] = new stwo_rocm_tree{trees[k]};
        }
        owner.trees.clear();  // the trees now belong to the caller's handles
    });
}

int stwo_rocm_gather_values(stwo_rocm_ctx* ctx, const uint32_t* const* cols, const uint32_t* log_sizes,
                            uint32_t n_cols, uint32_t words_per_value, uint32_t lifting_log,
                            const uint64_t* positions, uint32_t n_positions, uint32_t* out_host) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK((cols && log_sizes) || !n_cols, "null columns");
        SR_CHECK((positions && out_host) || !n_positions || !n_cols, "null argument");
        SR_CHECK(words_per_value == 1 || words_per_value == 4, "words per value must be 1 or 4");
        SR_CHECK(lifting_log <= 30, "lifting log out of range");
        if (!n_cols || !n_positions) return;
        std::vector<uint64_t> addr;
        addr.reserve((size_t)n_cols * n_positions * words_per_value);
        for (uint32_t i = 0; i < n_cols; i++) {
            SR_CHECK(log_sizes[i] >= 1 && log_sizes[i] <= lifting_log, "column log size out of range");
            const uint32_t sh = lifting_log - log_sizes[i] + 1;
            for (uint32_t q = 0; q < n_positions; q++) {
                SR_CHECK(positions[q] < (1ull << lifting_log), "position out of range");
                // lifted row (prover/vcs_lifted/leaves.zig:105-111)
                const uint64_t row = lifting_log == log_sizes[i] ? positions[q]
                                                                 : ((positions[q] >> sh) << 1) + (positions[q] & 1);
                const uint32_t* p = cols[i] + row * words_per_value;
                for (uint32_t w = 0; w < words_per_value; w++) addr.push_back((uint64_t)(uintptr_t)(p + w));
            }
        }
        DBuf da = upload(c, addr.data(), addr.size());
        DBuf o(&c, addr.size() * 4);
        launch_gather(c, da.as<uint64_t>(), addr.size(), o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(out_host, o.p, addr.size() * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_channel_create(stwo_rocm_ctx* ctx, const uint8_t digest[32], uint32_t n_draws,
                             stwo_rocm_channel** out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(out, "null output");
        DevChannel h;
        memset(&h, 0, sizeof h);
        if (digest) memcpy(h.digest, digest, 32);
        h.n_draws = n_draws;
        void* d = c.alloc(CH_BYTES);
        SR_HIP(hipMemcpyAsync(d, c.stage(&h, sizeof h), sizeof h, hipMemcpyHostToDevice, c.stream));
        c.sync();
        *out = new stwo_rocm_channel{(DevChannel*)d};
    });
}

int stwo_rocm_channel_destroy(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch) {
    return guarded(ctx, [&](Ctx& c) {
        if (!ch) return;
        c.sync();
        if (ch->d) c.free(ch->d);
        delete ch;
    });
}

int stwo_rocm_channel_state(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint8_t digest[32], uint32_t* n_draws) {
    return guarded(ctx, [&](Ctx& c) {
        DevChannel* d = chan_of(ch);
        DevChannel h;
        SR_HIP(hipMemcpyAsync(&h, d, sizeof h, hipMemcpyDeviceToHost, c.stream));
        c.sync();
        if (digest) memcpy(digest, h.digest, 32);
        if (n_draws) *n_draws = h.n_draws;
    });
}

int stwo_rocm_channel_mix_u32s(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* words, uint64_t n) {
    return guarded(ctx, [&](Ctx& c) { chan_mix_words(c, ch, words, n); });
}

int stwo_rocm_channel_mix_u64(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint64_t value) {
    return guarded(ctx, [&](Ctx& c) {
        const uint32_t w[2] = {(uint32_t)value, (uint32_t)(value >> 32)};
        chan_mix_words(c, ch, w, 2);
    });
}

int stwo_rocm_channel_mix_felts(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* felts, uint64_t n) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(felts || !n, "null felts");
        for (uint64_t i = 0; i < 4 * n; i++) SR_CHECK(felts[i] < P, "non-canonical M31 word");
        chan_mix_words(c, ch, felts, 4 * n);
    });
}

int stwo_rocm_channel_mix_root(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint8_t root[32]) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(root, "null root");
        uint32_t w[8];
        memcpy(w, root, 32);
        chan_mix_words(c, ch, w, 8);
    });
}

int stwo_rocm_channel_draw_secure_felt(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint32_t out[4]) {
    return guarded(ctx, [&](Ctx& c) {
        DevChannel* d = chan_of(ch);
        SR_CHECK(out, "null output");
        DBuf o(&c, 16);
        launch_tr_mix_draws(c, d, nullptr, 0, 1, o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(out, o.p, 16, hipMemcpyDeviceToHost, c.stream));
        chan_check(c, ch);
    });
}

int stwo_rocm_channel_draw_u32s(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint32_t n_blocks, uint32_t* out) {
    return guarded(ctx, [&](Ctx& c) {
        DevChannel* d = chan_of(ch);
        SR_CHECK(out || !n_blocks, "null output");
        SR_CHECK(n_blocks <= (1u << 20), "too many draws");
        if (!n_blocks) return;
        DBuf o(&c, 32ull * n_blocks);
        launch_tr_draw_words(c, d, n_blocks, o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(out, o.p, 32ull * n_blocks, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

}  // extern "C"
