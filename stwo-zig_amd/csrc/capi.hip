// C-ABI exports of the rocm backend (declared in include/stwo_rocm.h).
// Every entry point: lock the context, translate C++ exceptions into status
// codes, keep the message for stwo_rocm_last_error().
#include <string.h>

#include <algorithm>

#include "../../include/stwo_rocm.h"
#include "capi.hpp"

extern "C" {

int stwo_rocm_ctx_create(int device, uint64_t hbm_budget_bytes, stwo_rocm_ctx** out, char* err, size_t err_len) {
    if (!out) return STWO_ROCM_EINVAL;
    *out = nullptr;
    stwo_rocm_ctx* ctx = new stwo_rocm_ctx();
    try {
        ctx_init(ctx->c, device, hbm_budget_bytes);
        *out = ctx;
        return STWO_ROCM_OK;
    } catch (const Error& e) {
        if (err && err_len) {
            strncpy(err, e.what(), err_len - 1);
            err[err_len - 1] = 0;
        }
        ctx_release(ctx->c);
        delete ctx;
        return e.code;
    }
}

int stwo_rocm_ctx_destroy(stwo_rocm_ctx* ctx) {
    if (!ctx) return STWO_ROCM_EINVAL;
    {
        std::lock_guard<std::mutex> lk(ctx->c.mu);
        serve_pool_free(ctx->pool);
        ctx->pool = nullptr;
        ctx_release(ctx->c);
    }
    delete ctx;
    return STWO_ROCM_OK;
}

const char* stwo_rocm_last_error(stwo_rocm_ctx* ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

int stwo_rocm_ctx_identity(stwo_rocm_ctx* ctx, char* out, size_t out_len) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(out && out_len, "null output");
        hipDeviceProp_t p;
        SR_HIP(hipGetDeviceProperties(&p, c.device));
        std::string s = std::string(p.name) + ";" + p.gcnArchName + ";" + std::to_string(p.multiProcessorCount) +
                        ";" + std::to_string((unsigned long long)p.totalGlobalMem);
        strncpy(out, s.c_str(), out_len - 1);
        out[out_len - 1] = 0;
    });
}

int stwo_rocm_synchronize(stwo_rocm_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) { c.sync(); });
}

int stwo_rocm_memory_info(stwo_rocm_ctx* ctx, uint64_t* live, uint64_t* cached, uint64_t* peak) {
    return guarded(ctx, [&](Ctx& c) {
        if (live) *live = c.bytes_live;
        if (cached) *cached = c.bytes_cached;
        if (peak) *peak = c.peak_live;
    });
}

int stwo_rocm_release_cached(stwo_rocm_ctx* ctx) {
    return guarded(ctx, [&](Ctx& c) {
        c.trim();
        // the serving slots (their streams, arenas and towers) are rebuilt by the next serve call
        serve_pool_free(ctx->pool);
        ctx->pool = nullptr;
    });
}

int stwo_rocm_set_budget(stwo_rocm_ctx* ctx, uint64_t hbm_budget_bytes) {
    return guarded(ctx, [&](Ctx& c) {
        c.hbm_budget = hbm_budget_bytes;
        if (hbm_budget_bytes && c.bytes_live + c.bytes_cached > hbm_budget_bytes) c.trim();
    });
}

int stwo_rocm_get_runtime_info(stwo_rocm_ctx* ctx, int reset_peak, stwo_rocm_runtime_info* out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(out, "null output");
        memset(out, 0, sizeof(*out));
        out->tower_log = c.tower_log;
        out->tower_bytes = c.tower_log ? 4ull * 4ull * (1ull << (c.tower_log - 1)) : 0;
        out->tower_builds = c.tower_builds;
        out->bytes_live = c.bytes_live;
        out->bytes_cached = c.bytes_cached;
        out->peak_live = c.peak_live;
        if (reset_peak) c.peak_live = c.bytes_live;
    });
}

int stwo_rocm_job_device_bytes(const stwo_rocm_serve_job* job, uint64_t* bytes) {
    if (!job || !bytes) return STWO_ROCM_EINVAL;
    if (job->kind > 4 || job->log_n < 1 || job->log_n > 30 || (job->kind == 2 && job->log_n < 4)) return STWO_ROCM_EINVAL;
    *bytes = serve_job_bytes(ServeJob{job->kind, job->log_n, job->seq_len, 0, job->seed, 0.0});
    return STWO_ROCM_OK;
}

int stwo_rocm_buffer_alloc(stwo_rocm_ctx* ctx, uint64_t bytes, void** dptr) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(dptr, "null output");
        *dptr = c.alloc(bytes);
    });
}

int stwo_rocm_buffer_free(stwo_rocm_ctx* ctx, void* dptr) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(c.live_blocks.count(dptr), "pointer not owned by this context");
        c.free(dptr);
    });
}

int stwo_rocm_buffer_upload(stwo_rocm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&](Ctx& c) {
        if (!bytes) return;
        SR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c.stream));
        c.sync();
    });
}

int stwo_rocm_buffer_download(stwo_rocm_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
    return guarded(ctx, [&](Ctx& c) {
        if (!bytes) return;
        SR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_twiddles_prepare(stwo_rocm_ctx* ctx, uint32_t max_circle_log) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(max_circle_log >= 1 && max_circle_log <= 30, "circle log out of range");
        c.ensure_tower(max_circle_log);
        c.sync();
    });
}

int stwo_rocm_twiddles_read(stwo_rocm_ctx* ctx, uint32_t log, uint32_t* tw, uint32_t* itw) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(log >= 1 && log <= 30, "circle log out of range");
        c.ensure_tower(log);
        uint64_t n = 1ull << (log - 1);
        uint64_t off = (1ull << (c.tower_log - 1)) - n;  // suffix view (twiddle_tower.zig:66-87)
        if (tw) SR_HIP(hipMemcpyAsync(tw, c.d_tw + off, n * 4, hipMemcpyDeviceToHost, c.stream));
        if (itw) SR_HIP(hipMemcpyAsync(itw, c.d_itw + off, n * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_batch_inverse(stwo_rocm_ctx* ctx, const uint32_t* d_in, uint64_t n, uint32_t* d_out) {
    return guarded(ctx, [&](Ctx& c) {
        launch_batch_inverse(c, d_in, n, d_out);
        c.sync();
    });
}

int stwo_rocm_bit_reverse(stwo_rocm_ctx* ctx, uint32_t* d_vals, uint32_t log_size, uint32_t wpe) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(log_size <= 30 && (wpe == 1 || wpe == 4), "bit_reverse: bad shape");
        launch_bit_reverse(c, d_vals, log_size, wpe);
        c.sync();
    });
}

int stwo_rocm_accumulate(stwo_rocm_ctx* ctx, uint32_t* d_acc, const uint32_t* d_src, uint64_t n_words) {
    return guarded(ctx, [&](Ctx& c) {
        launch_accumulate(c, d_acc, d_src, n_words);
        c.sync();
    });
}

int stwo_rocm_circle_interpolate(stwo_rocm_ctx* ctx, uint32_t* const* cols, uint32_t n_cols, uint32_t log_size,
                                 double* gpu_ms) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(cols || !n_cols, "null columns");
        SR_CHECK(log_size >= 1 && log_size <= 30, "log size out of range");
        check_aligned((const void* const*)cols, n_cols, "circle_interpolate");
        if (log_size >= 3) c.ensure_tower(log_size);
        DBuf d = upload(c, cols, n_cols);
        double ms = timed(c, [&] {
            launch_circle_fft(c, true, d.as<const uint32_t* const>(), d.as<uint32_t* const>(), n_cols, log_size,
                              log_size);
        });
        if (gpu_ms) *gpu_ms = ms;
    });
}

int stwo_rocm_circle_evaluate(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                              uint32_t log_blowup, uint32_t* const* out, double* gpu_ms) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK((coeffs && out) || !n_cols, "null columns");
        check_aligned((const void* const*)coeffs, n_cols, "circle_evaluate");
        check_aligned((const void* const*)out, n_cols, "circle_evaluate");
        uint32_t m = log_size + log_blowup;
        SR_CHECK(m >= 1 && m <= 30, "log size out of range");
        if (m >= 3) c.ensure_tower(m);
        DBuf ds = upload(c, coeffs, n_cols), dd = upload(c, out, n_cols);
        double ms = timed(c, [&] {
            launch_circle_fft(c, false, ds.as<const uint32_t* const>(), dd.as<uint32_t* const>(), n_cols, m,
                              log_size);
        });
        if (gpu_ms) *gpu_ms = ms;
    });
}

int stwo_rocm_eval_at_point(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                            const uint32_t point[8], uint32_t* out_host) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(point && out_host && (coeffs || !n_cols), "null argument");
        SR_CHECK(log_size <= 30, "log size out of range");
        if (!n_cols) return;
        check_aligned((const void* const*)coeffs, n_cols, "eval_at_point");
        qm31 px = qin(point), py = qin(point + 4);
        qm31 f[32];
        if (log_size >= 1) f[0] = py;
        if (log_size >= 2) f[1] = px;
        for (uint32_t k = 2; k < log_size; k++) {
            qm31 s = qsqr(f[k - 1]);
            f[k] = qsub(qadd(s, s), qone());
        }
        DBuf w(&c, (1ull << log_size) * 16);
        launch_eval_weights(c, f, log_size, w.as<uint32_t>());
        DBuf d = upload(c, coeffs, n_cols);
        DBuf o(&c, (size_t)n_cols * 16);
        launch_eval_columns(c, d.as<const uint32_t* const>(), n_cols, log_size, w.as<uint32_t>(), o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(out_host, o.p, (size_t)n_cols * 16, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_merkle_commit(stwo_rocm_ctx* ctx, const uint32_t* const* cols, const uint32_t* log_sizes,
                            uint32_t n_cols, stwo_rocm_tree** out, uint8_t* root_out, double* gpu_ms) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(out, "null output");
        SR_CHECK((cols && log_sizes) || !n_cols, "null columns");
        std::vector<const uint32_t*> cv(cols, cols + n_cols);
        std::vector<uint32_t> lv(log_sizes, log_sizes + n_cols);
        DeviceTree* t = nullptr;
        double ms = timed(c, [&] { t = merkle_commit(c, cv, lv); });
        if (n_cols) tree_read_root(c, *t);
        if (root_out) memcpy(root_out, t->root, 32);
        *out = new stwo_rocm_tree{t};
        if (gpu_ms) *gpu_ms = ms;
    });
}

int stwo_rocm_merkle_parent_layer(stwo_rocm_ctx* ctx, const uint8_t* d_children, uint32_t log_size,
                                  uint8_t* d_parents) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(log_size <= 30, "log size out of range");
        launch_merkle_parent_layer(c, d_children, log_size, d_parents);
        c.sync();
    });
}

int stwo_rocm_tree_root(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, uint8_t root_out[32]) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(tree && root_out, "null argument");
        memcpy(root_out, tree->t->root, 32);
    });
}

int stwo_rocm_tree_max_log(stwo_rocm_tree* tree, uint32_t* out) {
    if (!tree || !out) return STWO_ROCM_EINVAL;
    *out = tree->t->max_log;
    return STWO_ROCM_OK;
}

int stwo_rocm_tree_read_hashes(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, uint32_t layer_log, const uint32_t* indices,
                               uint32_t n, uint8_t* out_host) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(tree && (indices || !n) && (out_host || !n), "null argument");
        SR_CHECK(layer_log <= tree->t->max_log, "layer out of range");
        if (!n) return;
        std::vector<uint64_t> addr;
        for (uint32_t i = 0; i < n; i++) {
            SR_CHECK(indices[i] < (1ull << layer_log), "hash index out of range");
            const uint32_t* p = (const uint32_t*)(tree->t->layer(layer_log) + (uint64_t)indices[i] * 32);
            for (int k = 0; k < 8; k++) addr.push_back((uint64_t)(uintptr_t)(p + k));
        }
        DBuf da = upload(c, addr.data(), addr.size());
        DBuf o(&c, addr.size() * 4);
        launch_gather(c, da.as<uint64_t>(), addr.size(), o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(out_host, o.p, addr.size() * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_tree_decommit(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, const uint64_t* positions, uint32_t n_positions,
                            uint8_t* witness_out, uint32_t witness_cap, uint32_t* n_witness) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(tree && n_witness && (positions || !n_positions), "null argument");
        const DeviceTree& t = *tree->t;
        std::vector<uint64_t> prev;
        for (uint32_t i = 0; i < n_positions; i++) {
            SR_CHECK(positions[i] < (1ull << t.max_log), "position out of range");
            SR_CHECK(i == 0 || positions[i - 1] <= positions[i], "positions must be sorted");
            if (i == 0 || positions[i - 1] != positions[i]) prev.push_back(positions[i]);
        }
        std::vector<uint64_t> addr;
        for (int32_t layer = (int32_t)t.max_log - 1; layer >= 0; layer--) {
            std::vector<uint64_t> cur;
            size_t q = 0;
            while (q < prev.size()) {
                uint64_t first = prev[q];
                bool sib = q + 1 < prev.size() && (first ^ 1) == prev[q + 1];
                if (!sib) {
                    const uint32_t* p = (const uint32_t*)(t.layer(layer + 1) + (first ^ 1) * 32);
                    for (int k = 0; k < 8; k++) addr.push_back((uint64_t)(uintptr_t)(p + k));
                }
                cur.push_back(first >> 1);
                q += sib ? 2 : 1;
            }
            prev.swap(cur);
        }
        *n_witness = (uint32_t)(addr.size() / 8);
        SR_CHECK(*n_witness <= witness_cap, "witness buffer too small");
        if (addr.empty()) return;
        SR_CHECK(witness_out, "null witness buffer");
        DBuf da = upload(c, addr.data(), addr.size());
        DBuf o(&c, addr.size() * 4);
        launch_gather(c, da.as<uint64_t>(), addr.size(), o.as<uint32_t>());
        SR_HIP(hipMemcpyAsync(witness_out, o.p, addr.size() * 4, hipMemcpyDeviceToHost, c.stream));
        c.sync();
    });
}

int stwo_rocm_tree_free(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree) {
    return guarded(ctx, [&](Ctx& c) {
        if (!tree) return;
        c.sync();
        tree_free(c, tree->t);
        delete tree;
    });
}

int stwo_rocm_fri_fold_line(stwo_rocm_ctx* ctx, const uint32_t* d_src, uint32_t log_src, const uint32_t alpha[4],
                            uint32_t* d_dst) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_src && d_dst && alpha, "null argument");
        SR_CHECK(log_src >= 1 && log_src <= 29, "log size out of range");
        launch_fold_line(c, d_src, log_src, qin(alpha), d_dst, nullptr);
        c.sync();
    });
}

int stwo_rocm_fri_fold_circle_into_line(stwo_rocm_ctx* ctx, const uint32_t* const* src_coords, uint32_t log_src,
                                        const uint32_t alpha[4], uint32_t* d_dst) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(src_coords && d_dst && alpha, "null argument");
        SR_CHECK(log_src >= 1 && log_src <= 30, "log size out of range");
        launch_fold_circle_into_line(c, src_coords, log_src, qin(alpha), d_dst, true, nullptr);
        c.sync();
    });
}

int stwo_rocm_compute_quotients(stwo_rocm_ctx* ctx, uint32_t L, const uint32_t* const* columns,
                                const uint32_t* column_log_sizes, uint32_t n_columns, const uint32_t* sample_columns,
                                const uint32_t* sample_points, const uint32_t* sample_values, uint32_t n_samples,
                                const uint32_t random_coeff[4], uint32_t* const* out_coords,
                                stwo_rocm_tree** tree_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(L >= 1 && L <= 30, "lifting log size out of range");
        SR_CHECK(out_coords && random_coeff, "null argument");
        SR_CHECK((columns && column_log_sizes) || !n_columns, "null columns");
        SR_CHECK((sample_columns && sample_points && sample_values) || !n_samples, "null samples");
        for (uint32_t i = 0; i < n_columns; i++)
            SR_CHECK(column_log_sizes[i] >= 1 && column_log_sizes[i] <= L, "column log size out of range");
        // Samples per column in input order (tree/column/sample order).
        std::vector<std::vector<ColumnSample>> per_col(n_columns);
        for (uint32_t s = 0; s < n_samples; s++) {
            SR_CHECK(sample_columns[s] < n_columns, "sample column out of range");
            SR_CHECK(s == 0 || sample_columns[s - 1] <= sample_columns[s], "samples must be in column order");
            per_col[sample_columns[s]].push_back(
                {{qin(sample_points + 8 * s), qin(sample_points + 8 * s + 4)}, qin(sample_values + 4 * s)});
        }
        std::vector<uint32_t> logs(column_log_sizes, column_log_sizes + n_columns);
        QuotientPlan qp = plan_quotients(c, L, logs, per_col, qin(random_coeff));
        std::vector<uint32_t>& ecols = qp.entry_cols;
        std::vector<uint32_t>& ecf = qp.entry_coeffs;
        std::vector<QuotientBatchDev>& batches = qp.batches;
        DeviceTree* t = nullptr;
        try {
            if (tree_out) {
                t = new DeviceTree();
                t->max_log = L;
                t->layers = (uint8_t*)c.alloc(((2ull << L) - 1) * 32);
            }
            launch_quotients(c, columns, column_log_sizes, ecols.data(), ecf.data(), (uint32_t)ecols.size(),
                             batches.data(), (uint32_t)batches.size(), L, out_coords, t ? t->layer(L) : nullptr);
            if (t) {
                launch_merkle_parents(c, *t);
                tree_read_root(c, *t);
                *tree_out = new stwo_rocm_tree{t};
            }
            c.sync();
        } catch (...) {
            if (t) tree_free(c, t);
            throw;
        }
    });
}

int stwo_rocm_widefib_trace(stwo_rocm_ctx* ctx, uint32_t log_n_rows, uint32_t seq_len, uint32_t mode, uint64_t seed,
                            uint32_t* d_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_out, "null output");
        SR_CHECK(log_n_rows >= 1 && log_n_rows <= 30 && seq_len >= 2 && mode <= 1, "bad trace shape");
        launch_widefib_trace(c, log_n_rows, seq_len, mode, seed, d_out, 1ull << log_n_rows);
        c.sync();
    });
}

int stwo_rocm_widefib_composition(stwo_rocm_ctx* ctx, const uint32_t* const* eval_cols, uint32_t n_cols,
                                  uint32_t log_n_rows, const uint32_t random_coeff[4], uint32_t* const* out_coords) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(eval_cols && out_coords && random_coeff, "null argument");
        SR_CHECK(n_cols >= 3 && log_n_rows >= 1 && log_n_rows <= 29, "bad shape");
        qm31 a = qin(random_coeff);
        uint32_t NC = n_cols - 2, CL = log_n_rows + 1;
        std::vector<uint32_t> pw(4 * NC);
        qm31 cur = qone();
        for (uint32_t k = 0; k < NC; k++) {
            qstore(cur, &pw[4 * k]);
            cur = qmul(cur, a);
        }
        uint32_t den[2];
        for (int r = 0; r < 2; r++) {
            pt dp = idx_to_point(c.gen, domain_index_at(CL, (uint64_t)r));
            pt init = idx_to_point(c.gen, subgroup_gen(log_n_rows + 1));
            pt s = padd(padd(dp, pconj(init)), init);
            uint32_t x = s.x;
            for (uint32_t i = 1; i < log_n_rows; i++) x = mdouble_x(x);
            den[r] = minv(x);
        }
        DBuf dp = upload(c, pw.data(), pw.size());
        DBuf dc = upload(c, eval_cols, n_cols);
        launch_widefib_composition(c, dc.as<const uint32_t* const>(), n_cols, CL, log_n_rows, dp.as<uint32_t>(),
                                   den[0], den[1], out_coords);
        c.sync();
    });
}

int stwo_rocm_pow_grind(stwo_rocm_ctx* ctx, const uint8_t digest[32], uint32_t pow_bits, uint64_t* nonce) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(digest && nonce, "null argument");
        SR_CHECK(pow_bits <= 40, "pow_bits out of range");
        Channel ch;
        memcpy(ch.digest, digest, 32);
        *nonce = 0;
        if (pow_bits == 0) return;
        uint8_t prefix[32];
        ch.pow_prefix(pow_bits, prefix);
        uint64_t chunk = pow_bits <= 16 ? (1ull << 20) : (1ull << 24);
        for (uint64_t base = 0;; base += chunk) {
            uint64_t r = launch_pow_grind(c, prefix, pow_bits, base, chunk);
            if (r != UINT64_MAX) {
                *nonce = r;
                return;
            }
        }
    });
}

// Runs the proof core and copies the proof and the optional stage timings out.
static void prove_out(Ctx& c, const ProveParams& p, const uint32_t* d_trace, uint8_t* proof_out, uint64_t cap,
                      uint64_t* proof_len, stwo_rocm_prove_stats* stats) {
    std::string proof;
    ProveStats st;
    prove_wide_fibonacci(c, p, d_trace, proof, stats ? &st : nullptr);
    *proof_len = proof.size();
    if (stats) {
        for (int i = 0; i < STAGES; i++) {
            stats->stage_gpu_ms[i] = st.stage_gpu_ms[i];
            stats->stage_host_ms[i] = st.stage_host_ms[i];
        }
        stats->total_gpu_ms = st.total_gpu_ms;
        stats->total_host_ms = st.total_host_ms;
        stats->peak_hbm_bytes = st.peak_hbm_bytes;
    }
    SR_CHECK(proof.size() <= cap, "proof buffer too small");
    SR_CHECK(proof_out, "null proof buffer");
    memcpy(proof_out, proof.data(), proof.size());
}

int stwo_rocm_prove_wide_fibonacci(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params,
                                   const uint32_t* d_trace, uint8_t* proof_out, uint64_t cap, uint64_t* proof_len,
                                   stwo_rocm_prove_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(params && proof_len, "null argument");
        ProveParams p;
        p.log_n_rows = params->log_n_rows;
        p.sequence_len = params->sequence_len;
        p.pow_bits = params->pow_bits;
        p.log_blowup_factor = params->log_blowup_factor;
        p.log_last_layer_degree_bound = params->log_last_layer_degree_bound;
        p.n_queries = params->n_queries;
        p.trace_mode = params->trace_mode;
        p.trace_seed = params->trace_seed;
        SR_CHECK(p.trace_mode <= 1, "trace_mode must be 0 or 1");
        prove_out(c, p, d_trace, proof_out, cap, proof_len, stats);
    });
}

static ProveParams batch_params(const stwo_rocm_widefib_params* params) {
    SR_CHECK(params, "null argument");
    ProveParams p;
    p.log_n_rows = params->log_n_rows;
    p.sequence_len = params->sequence_len;
    p.pow_bits = params->pow_bits;
    p.log_blowup_factor = params->log_blowup_factor;
    p.log_last_layer_degree_bound = params->log_last_layer_degree_bound;
    p.n_queries = params->n_queries;
    p.trace_mode = params->trace_mode;
    p.trace_seed = params->trace_seed;
    SR_CHECK(p.trace_mode <= 1, "trace_mode must be 0 or 1");
    return p;
}

// Offsets, stats and the proofs back to back (the copy on host threads).
static void batch_out(std::vector<std::string>& proofs, const BatchStats& st, uint8_t* proofs_out,
                      uint64_t cap, uint64_t* offsets, stwo_rocm_batch_stats* stats) {
    const uint32_t n = (uint32_t)proofs.size();
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; k++) {
        offsets[k] = total;
        total += proofs[k].size();
    }
    offsets[n] = total;
    if (stats) {
        stats->device_ms = st.device_ms;
        stats->tail_ms = st.tail_ms;
        stats->total_ms = st.total_ms;
        stats->slice_bytes = st.slice_bytes;
        for (int i = 0; i < 6; i++) stats->phase_ms[i] = st.phase_ms[i];
    }
    SR_CHECK(proofs_out && total <= cap, "proof buffer too small");
    parallel_for(n, [&](uint32_t k) {
        memcpy(proofs_out + offsets[k], proofs[k].data(), proofs[k].size());
        std::string().swap(proofs[k]);  // freed on the pool
    });
}

int stwo_rocm_prove_wide_fibonacci_batch(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params,
                                         uint32_t n_proofs, const uint64_t* trace_seeds, uint8_t* proofs_out,
                                         uint64_t cap, uint64_t* offsets, stwo_rocm_batch_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(offsets, "null argument");
        const ProveParams p = batch_params(params);
        std::vector<std::string> proofs;
        BatchStats st;
        prove_wide_fibonacci_batch(c, p, n_proofs, trace_seeds, proofs, &st);
        batch_out(proofs, st, proofs_out, cap, offsets, stats);
    });
}

int stwo_rocm_batch_prover_create(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params, uint32_t n_proofs,
                                  stwo_rocm_batch_prover** out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(out, "null argument");
        *out = nullptr;
        const ProveParams p = batch_params(params);
        std::unique_ptr<stwo_rocm_batch_prover> h(new stwo_rocm_batch_prover());
        h->p = batch_plan_create(c, p, n_proofs);
        h->n = n_proofs;
        *out = h.release();
    });
}

int stwo_rocm_batch_prover_run(stwo_rocm_ctx* ctx, stwo_rocm_batch_prover* bp, const uint64_t* trace_seeds,
                               int use_graph, uint8_t* proofs_out, uint64_t cap, uint64_t* offsets,
                               stwo_rocm_batch_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(bp && bp->p && offsets, "null argument");
        std::vector<std::string> proofs;
        BatchStats st;
        batch_plan_run(c, *bp->p, trace_seeds, use_graph != 0, proofs, &st);
        batch_out(proofs, st, proofs_out, cap, offsets, stats);
    });
}

int stwo_rocm_batch_prover_destroy(stwo_rocm_ctx* ctx, stwo_rocm_batch_prover* bp) {
    if (!bp) return STWO_ROCM_OK;
    const int rc = guarded(ctx, [&](Ctx& c) {
        c.sync();
        batch_plan_free(bp->p);
        bp->p = nullptr;
    });
    delete bp;
    return rc;
}

int stwo_rocm_prove_poseidon(stwo_rocm_ctx* ctx, const stwo_rocm_poseidon_params* params, uint8_t* proof_out,
                             uint64_t cap, uint64_t* proof_len, stwo_rocm_prove_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(params && proof_len, "null argument");
        SR_CHECK(params->log_n_instances >= 4 && params->log_n_instances <= 29, "log_n_instances must be 4..29");
        ProveParams p;
        p.log_n_rows = params->log_n_instances - 3;
        p.sequence_len = POSEIDON_COLUMNS;
        p.pow_bits = params->pow_bits;
        p.log_blowup_factor = params->log_blowup_factor;
        p.log_last_layer_degree_bound = params->log_last_layer_degree_bound;
        p.n_queries = params->n_queries;
        p.trace_mode = 0;
        p.trace_seed = 0;
        p.air = AIR_POSEIDON;
        prove_out(c, p, nullptr, proof_out, cap, proof_len, stats);
    });
}

int stwo_rocm_prove_xor(stwo_rocm_ctx* ctx, const stwo_rocm_xor_params* params, uint8_t* proof_out, uint64_t cap,
                        uint64_t* proof_len, stwo_rocm_prove_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(params && proof_len, "null argument");
        SR_CHECK(params->log_size >= 1 && params->log_size <= 26, "log_size must be 1..26");
        SR_CHECK(params->log_step <= params->log_size, "log_step must not exceed log_size");
        ProveParams p;
        p.log_n_rows = params->log_size;
        p.sequence_len = 1;
        p.pow_bits = params->pow_bits;
        p.log_blowup_factor = params->log_blowup_factor;
        p.log_last_layer_degree_bound = params->log_last_layer_degree_bound;
        p.n_queries = params->n_queries;
        p.trace_mode = 0;
        p.trace_seed = 0;
        p.air = AIR_XOR;
        p.xor_log_step = params->log_step;
        p.xor_offset = params->offset;
        prove_out(c, p, nullptr, proof_out, cap, proof_len, stats);
    });
}

// ---------------------------------------------------------------- LogUp / GKR
int stwo_rocm_prove_logup_xor(stwo_rocm_ctx* ctx, const stwo_rocm_logup_params* params, uint8_t* proof_out,
                              uint64_t cap, uint64_t* proof_len) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(params && proof_len, "null argument");
        ProveParams p{};
        p.pow_bits = params->pow_bits;
        p.log_blowup_factor = params->log_blowup_factor;
        p.log_last_layer_degree_bound = params->log_last_layer_degree_bound;
        p.n_queries = params->n_queries;
        std::string proof;
        prove_logup_xor(c, params->log_n, params->seed, p, proof);
        *proof_len = proof.size();
        SR_CHECK(proof_out && cap >= proof.size(), "proof buffer too small");
        memcpy(proof_out, proof.data(), proof.size());
    });
}

int stwo_rocm_xor_trace(stwo_rocm_ctx* ctx, uint32_t log_size, uint64_t seed, uint32_t* d_a, uint32_t* d_b,
                        uint32_t* d_c, uint32_t* d_mult) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_a && d_b && d_c && d_mult, "null argument");
        launch_xor_trace(c, log_size, seed, d_a, d_b, d_c, d_mult);
        c.sync();
    });
}

int stwo_rocm_logup_denominators(stwo_rocm_ctx* ctx, const uint32_t* d_a, const uint32_t* d_b, const uint32_t* d_c,
                                 uint64_t n, const uint32_t z[4], const uint32_t alpha[4], uint32_t* d_den) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(z && alpha && d_den && n > 0, "null argument");
        SR_CHECK((d_a && d_b && d_c) || (!d_a && !d_b && !d_c), "lookup columns: all or none");
        launch_logup_denominators(c, d_a, d_b, d_c, n, qload(z), qload(alpha), d_den);
        c.sync();
    });
}

int stwo_rocm_logup_cumulative(stwo_rocm_ctx* ctx, int num_kind, const uint32_t* d_num, const uint32_t* d_den,
                               uint64_t n, uint32_t* d_sums, uint32_t claimed[4]) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_den && d_sums && claimed, "null argument");
        qstore(logup_cumulative(c, num_kind, d_num, d_den, n, d_sums), claimed);
    });
}

int stwo_rocm_logup_cumulative_noalloc(stwo_rocm_ctx* ctx, int num_kind, const uint32_t* d_num,
                                       const uint32_t* d_den, uint64_t n, uint32_t* d_sums, uint32_t claimed[4]) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_den && d_sums && claimed, "null argument");
        qstore(logup_cumulative_noalloc(c, num_kind, d_num, d_den, n, d_sums), claimed);
    });
}

int stwo_rocm_gkr_gen_eq_evals(stwo_rocm_ctx* ctx, const uint32_t* y, uint32_t n, const uint32_t scale[4],
                               uint32_t* d_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(scale && d_out && (y || n == 0), "null argument");
        DBuf dy = upload(c, y ? y : scale, y ? 4 * (size_t)n : 4);
        launch_gkr_eq_evals(c, dy.as<uint32_t>(), n, qload(scale), d_out);
        c.sync();
    });
}

int stwo_rocm_gkr_next_layer(stwo_rocm_ctx* ctx, int kind, const uint32_t* d_num, const uint32_t* d_den,
                             uint32_t n_vars, uint32_t* d_out_num, uint32_t* d_out_den) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_den && d_out_den, "null argument");
        launch_gkr_next_layer(c, kind, d_num, d_den, n_vars, d_out_num, d_out_den);
        c.sync();
    });
}

int stwo_rocm_gkr_sum_as_poly_in_first_variable(stwo_rocm_ctx* ctx, int kind, const uint32_t* d_eq,
                                                const uint32_t* d_num, const uint32_t* d_den, uint32_t n_vars,
                                                const uint32_t lambda[4], const uint32_t claim[4], const uint32_t* y,
                                                uint32_t y_len, const uint32_t correction[4], uint32_t coeffs_out[16],
                                                uint32_t* n_coeffs) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_eq && d_den && lambda && claim && y && correction && coeffs_out && n_coeffs, "null argument");
        SR_CHECK(kind >= 0 && kind <= 3, "gkr: layer kind");
        SR_CHECK(kind == GKR_GRAND_PRODUCT || kind == GKR_LOGUP_SINGLES || d_num, "gkr: numerators");
        std::vector<qm31> yy(y_len);
        for (uint32_t i = 0; i < y_len; i++) yy[i] = qload(y + 4 * i);
        std::vector<qm31> cf = gkr_sum_as_poly(c, kind, d_eq, d_num, d_den, n_vars, qload(lambda), qload(claim), yy,
                                               qload(correction));
        SR_CHECK(cf.size() <= 4, "gkr: round polynomial degree");
        *n_coeffs = (uint32_t)cf.size();
        for (size_t i = 0; i < cf.size(); i++) qstore(cf[i], coeffs_out + 4 * i);
    });
}

int stwo_rocm_mle_fix_first_variable(stwo_rocm_ctx* ctx, int base_field, const uint32_t* d_in, uint32_t n_vars,
                                     const uint32_t r[4], uint32_t* d_out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(d_in && r && d_out, "null argument");
        launch_mle_fix_first(c, base_field != 0, d_in, n_vars, qload(r), d_out);
        c.sync();
    });
}

int stwo_rocm_gkr_prove_batch(stwo_rocm_ctx* ctx, uint32_t n_instances, const int* kinds,
                              const uint32_t* const* d_nums, const uint32_t* const* d_dens, const uint32_t* n_vars,
                              const uint8_t channel_digest[32], uint32_t* words_out, uint64_t cap,
                              uint64_t* n_words) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(kinds && d_nums && d_dens && n_vars && channel_digest && n_words && n_instances > 0,
                 "null argument");
        std::vector<GkrInstance> inst(n_instances);
        for (uint32_t k = 0; k < n_instances; k++) inst[k] = {kinds[k], d_nums[k], d_dens[k], n_vars[k]};
        Channel ch;
        memcpy(ch.digest, channel_digest, 32);
        std::vector<uint32_t> w = gkr_prove_batch(c, ch, inst);
        *n_words = w.size();
        SR_CHECK(words_out && w.size() <= cap, "gkr: output buffer too small");
        memcpy(words_out, w.data(), 4 * w.size());
    });
}

int stwo_rocm_logup_xor_pipeline(stwo_rocm_ctx* ctx, uint32_t log_n, uint64_t seed, uint32_t* words_out,
                                 uint64_t cap, uint64_t* n_words, double* total_ms) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(n_words, "null argument");
        std::vector<uint32_t> w = logup_xor_pipeline(c, log_n, seed, total_ms);
        *n_words = w.size();
        SR_CHECK(words_out && w.size() <= cap, "logup: output buffer too small");
        memcpy(words_out, w.data(), 4 * w.size());
    });
}

// ---------------------------------------------------------------- batched pipeline
struct stwo_rocm_batched {
    BatchedPipeline* p;
};

int stwo_rocm_batched_create(stwo_rocm_ctx* ctx, const stwo_rocm_batched_spec* spec, stwo_rocm_batched** out) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(spec && out, "null argument");
        BatchedSpec s{spec->n_columns, spec->log_min, spec->log_max, spec->log_blowup, spec->fri_log, spec->n_folds,
                      spec->seed};
        auto* b = new stwo_rocm_batched{nullptr};
        try {
            b->p = new BatchedPipeline(c, s);
            b->p->capture();
        } catch (...) {
            delete b->p;
            delete b;
            throw;
        }
        *out = b;
    });
}

int stwo_rocm_batched_set_tuning(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t fft_tile_log,
                                 uint32_t merkle_levels) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(b && b->p, "null argument");
        SR_CHECK(fft_tile_log >= 11 && fft_tile_log <= 13, "fft_tile_log must be 11..13");
        SR_CHECK(merkle_levels >= 2 && merkle_levels <= 10, "merkle_levels must be 2..10");
        (void)c;
        b->p->set_tuning(fft_tile_log, merkle_levels);
    });
}

int stwo_rocm_batched_autotune(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t reps, uint32_t* fft_tile_log,
                               uint32_t* merkle_levels, char* json_out, size_t cap) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(b && b->p && fft_tile_log && merkle_levels, "null argument");
        (void)c;
        b->p->autotune(reps ? reps : 3);
        *fft_tile_log = b->p->tile_log;
        *merkle_levels = b->p->levels;
        if (json_out && cap) {
            std::string js = "[";
            for (size_t i = 0; i < b->p->tuning.size(); i++) {
                char buf[128];
                snprintf(buf, sizeof buf, "%s{\"fft_tile_log\":%u,\"merkle_levels\":%u,\"ms\":%.4f}", i ? "," : "",
                         b->p->tuning[i].fft_tile_log, b->p->tuning[i].merkle_levels, b->p->tuning[i].ms);
                js += buf;
            }
            js += "]";
            SR_CHECK(js.size() < cap, "autotune: json buffer too small");
            memcpy(json_out, js.c_str(), js.size() + 1);
        }
    });
}

int stwo_rocm_batched_run(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t n_runs, int use_graph,
                          double* ms_per_run, uint32_t* words_out, uint64_t cap, uint64_t* n_words) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(b && b->p && n_words, "null argument");
        double ms = 0;
        if (use_graph) {
            ms = b->p->replay(n_runs);
        } else {
            hipEvent_t e0, e1;
            SR_HIP(hipEventCreate(&e0));
            SR_HIP(hipEventCreate(&e1));
            SR_HIP(hipEventRecord(e0, c.stream));
            for (uint32_t i = 0; i < n_runs; i++) b->p->record();
            SR_HIP(hipEventRecord(e1, c.stream));
            SR_HIP(hipEventSynchronize(e1));
            float f = 0;
            SR_HIP(hipEventElapsedTime(&f, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            ms = n_runs ? f / n_runs : 0.0;
        }
        if (ms_per_run) *ms_per_run = ms;
        std::vector<uint32_t> w = b->p->result();
        *n_words = w.size();
        SR_CHECK(words_out && w.size() <= cap, "batched: output buffer too small");
        memcpy(words_out, w.data(), 4 * w.size());
    });
}

int stwo_rocm_batched_destroy(stwo_rocm_ctx* ctx, stwo_rocm_batched* b) {
    return guarded(ctx, [&](Ctx& c) {
        (void)c;
        if (!b) return;
        c.sync();
        delete b->p;
        delete b;
    });
}

// ---------------------------------------------------------------- serving
int stwo_rocm_serve(stwo_rocm_ctx* ctx, const stwo_rocm_serve_job* jobs, uint32_t n_jobs, uint32_t n_slots,
                    uint64_t hbm_budget_bytes, stwo_rocm_serve_result* results, stwo_rocm_serve_stats* stats) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(jobs && results && stats && n_jobs > 0, "null argument");
        std::vector<ServeJob> js(n_jobs);
        for (uint32_t i = 0; i < n_jobs; i++)
            js[i] = ServeJob{jobs[i].kind, jobs[i].log_n, jobs[i].seq_len, 0, jobs[i].seed, jobs[i].arrival_ms};
        std::vector<ServeOutcome> out;
        ServeStats st{};
        if (!ctx->pool) ctx->pool = serve_pool_new();
        serve(*ctx->pool, c.device, js, n_slots, hbm_budget_bytes, out, st);
        for (uint32_t i = 0; i < n_jobs; i++) {
            memcpy(results[i].digest, out[i].digest, 32);
            results[i].arrival_ms = out[i].arrival_ms;
            results[i].start_ms = out[i].start_ms;
            results[i].end_ms = out[i].end_ms;
            results[i].slot = out[i].slot;
            results[i].status = out[i].status;
            results[i].output_bytes = out[i].output_bytes;
        }
        stats->makespan_ms = st.makespan_ms;
        stats->p50_ms = st.p50_ms;
        stats->p99_ms = st.p99_ms;
        stats->mean_ms = st.mean_ms;
        stats->jobs_per_s = st.jobs_per_s;
        stats->peak_admitted_bytes = st.peak_admitted_bytes;
        stats->peak_device_bytes = st.peak_device_bytes;
        stats->max_in_flight = st.max_in_flight;
        stats->n_slots = st.n_slots;
    });
}

int stwo_rocm_profile_enable(stwo_rocm_ctx* ctx, int enable) {
    return guarded(ctx, [&](Ctx& c) {
        c.sync();
        c.profiling = enable != 0;
        c.krecs.clear();
        c.ev_used = 0;
    });
}

int stwo_rocm_profile_select(stwo_rocm_ctx* ctx, const char* kernel_class) {
    return guarded(ctx, [&](Ctx& c) {
        c.profile_only.clear();
        std::string all = kernel_class ? kernel_class : "";
        size_t at = 0;
        while (at < all.size()) {
            size_t bar = all.find('|', at);
            if (bar == std::string::npos) bar = all.size();
            if (bar > at) c.profile_only.push_back(all.substr(at, bar - at));
            at = bar + 1;
        }
    });
}

int stwo_rocm_profile_sample(stwo_rocm_ctx* ctx, uint32_t every) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(every >= 1, "profile_sample: every must be >= 1");
        c.profile_every = every;
        c.profile_seen = 0;
    });
}

int stwo_rocm_profile_read(stwo_rocm_ctx* ctx, char* json_out, size_t cap, size_t* len) {
    return guarded(ctx, [&](Ctx& c) {
        SR_CHECK(len, "null argument");
        c.sync();
        struct Agg {
            uint64_t launches = 0;
            double ms = 0, bytes = 0, max_ms = 0;
        };
        std::map<std::string, Agg> agg;
        for (auto& r : c.krecs) {
            float ms = 0;
            SR_HIP(hipEventElapsedTime(&ms, r.a, r.b));
            Agg& a = agg[r.cls];
            a.launches++;
            a.ms += ms;
            a.bytes += r.bytes;
            if (ms > a.max_ms) a.max_ms = ms;
        }
        std::string s = "{";
        bool first = true;
        for (auto& kv : agg) {
            if (!first) s += ",";
            first = false;
            char buf[256];
            snprintf(buf, sizeof buf, "\"%s\":{\"launches\":%llu,\"ms\":%.6f,\"max_ms\":%.6f,\"bytes\":%.0f}",
                     kv.first.c_str(), (unsigned long long)kv.second.launches, kv.second.ms, kv.second.max_ms,
                     kv.second.bytes);
            s += buf;
        }
        s += "}";
        *len = s.size();
        SR_CHECK(json_out && s.size() < cap, "profile buffer too small");
        memcpy(json_out, s.c_str(), s.size() + 1);
        c.krecs.clear();
        c.ev_used = 0;
    });
}

}  // extern "C"
