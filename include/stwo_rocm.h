/*
 * stwo_rocm.h -- C-ABI of the MI355X-native (gfx950) proving backend for stwo-zig.
 *
 * This is the drop-in boundary a Zig `RocmBackend` binds with `extern fn`, in the
 * place of the Metal runtime ABI (src/backends/metal/runtime/bindings.zig:11-654)
 * and of the external CUDA library's FFI (src/backends/cuda/ffi.zig:27-280).
 *
 * Conventions (modelled on the Metal ABI, not on the CUDA one, which returns void):
 *   - every entry point returns an int status: STWO_ROCM_OK (0) or an error code;
 *     the message of the last failure is available from stwo_rocm_last_error();
 *   - sizes are uint64_t / uint32_t log sizes; no C++ or torch types cross the ABI;
 *   - "device pointers" are HBM addresses obtained from stwo_rocm_buffer_alloc;
 *     arrays *of* device pointers (`cols`) are host arrays, as in ffi.zig
 *     (`values: [*][*]u32`); every device pointer must be 16-byte aligned (columns and
 *     QM31 arrays are read with 16-byte loads; buffer_alloc returns 256-byte aligned
 *     memory, and a column inside a buffer starts at a multiple of 4 words); the column
 *     entries check this and fail with STWO_ROCM_EINVAL;
 *   - M31 values are canonical u32 words in [0, 2^31-1); a QM31 is 4 words
 *     (c0.a, c0.b, c1.a, c1.b) (src/core/fields/qm31.zig:64-70); a circle point over
 *     QM31 is 8 words (x then y); hashes are 32 bytes;
 *   - columns are stored in bit-reversed circle-domain order, as in the reference;
 *   - calls on one context are serialised by an internal mutex and execute on the
 *     context's stream; results are complete when a call returns unless stated.
 */
#ifndef STWO_ROCM_H
#define STWO_ROCM_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define STWO_ROCM_OK 0
#define STWO_ROCM_EINVAL 1     /* invalid argument / shape */
#define STWO_ROCM_ENOMEM 2     /* device allocation failed */
#define STWO_ROCM_EDEVICE 3    /* HIP runtime error */
#define STWO_ROCM_EBUDGET 4    /* HBM budget of the context exceeded */
#define STWO_ROCM_EINTERNAL 5  /* anything else (e.g. constraints not satisfied) */

typedef struct stwo_rocm_ctx stwo_rocm_ctx;
typedef struct stwo_rocm_tree stwo_rocm_tree;

/* ------------------------------------------------------------------ runtime */
/* Replaces the Metal shared runtime init (src/backends/metal/shared_runtime.zig:1-40;
 * MetalCommitBackend.initializeRuntime, metal/commit_backend.zig). hbm_budget_bytes = 0:
 * unlimited.  On failure *out is NULL and err (if non-NULL) receives the message. */
int stwo_rocm_ctx_create(int device, uint64_t hbm_budget_bytes, stwo_rocm_ctx** out, char* err, size_t err_len);
int stwo_rocm_ctx_destroy(stwo_rocm_ctx* ctx);
/* Changes the context's device budget (0 = unlimited) for later allocations, like the
 * reference's per-request budget_gib (src/tools/metal_session/protocol.zig:51); when the bytes
 * live and cached exceed it the cached arena blocks are released. */
int stwo_rocm_set_budget(stwo_rocm_ctx* ctx, uint64_t hbm_budget_bytes);
const char* stwo_rocm_last_error(stwo_rocm_ctx* ctx);
/* runtimePlatformIdentityAlloc (metal/commit_backend.zig): "name;arch;CUs;HBM bytes". */
int stwo_rocm_ctx_identity(stwo_rocm_ctx* ctx, char* out, size_t out_len);
int stwo_rocm_synchronize(stwo_rocm_ctx* ctx);
/* telemetrySnapshot (metal/telemetry.zig): arena bytes live / cached / peak.  release_cached
 * frees the cached arena blocks and the serving slots that stwo_rocm_serve keeps resident
 * between calls (the next serve call rebuilds them). */
int stwo_rocm_memory_info(stwo_rocm_ctx* ctx, uint64_t* live, uint64_t* cached, uint64_t* peak);
int stwo_rocm_release_cached(stwo_rocm_ctx* ctx);
/* Session fields of the reference report (src/prover/native/report.zig:73-78: max circle log,
 * retained twiddle bytes, tower build count) for the device context, with the arena counters.
 * reset_peak != 0 restarts the peak at the bytes live now (after reading it). */
typedef struct {
    uint32_t tower_log; /* circle log of the resident twiddle tower; 0 before the first build */
    uint32_t reserved;
    uint64_t tower_bytes;  /* device bytes of the tower (twiddles, inverses and their doubled copies) */
    uint64_t tower_builds; /* builds of the tower since the context was created */
    uint64_t bytes_live, bytes_cached, peak_live;
} stwo_rocm_runtime_info;
int stwo_rocm_get_runtime_info(stwo_rocm_ctx* ctx, int reset_peak, stwo_rocm_runtime_info* out);

/* ------------------------------------------------------------------ buffers */
/* cuda_malloc_uint32_t / cuda_free_memory / copy_uint32_t_vec_* (cuda/ffi.zig:27-37,82). */
int stwo_rocm_buffer_alloc(stwo_rocm_ctx* ctx, uint64_t bytes, void** dptr);
int stwo_rocm_buffer_free(stwo_rocm_ctx* ctx, void* dptr);
int stwo_rocm_buffer_upload(stwo_rocm_ctx* ctx, void* dst_dev, const void* src_host, uint64_t bytes);
int stwo_rocm_buffer_download(stwo_rocm_ctx* ctx, void* dst_host, const void* src_dev, uint64_t bytes);

/* ------------------------------------------------------------------ twiddles */
/* precompute_twiddles (cuda/ffi.zig:63); M31TwiddleTower.init (prover/poly/twiddle_tower.zig:29-64).
 * Makes the resident tower cover circle domains up to max_circle_log. */
int stwo_rocm_twiddles_prepare(stwo_rocm_ctx* ctx, uint32_t max_circle_log);
/* M31TwiddleTower.view (twiddle_tower.zig:66-87): copies the 2^(log-1) twiddles and
 * inverse twiddles of CanonicCoset(circle_log) to host memory. */
int stwo_rocm_twiddles_read(stwo_rocm_ctx* ctx, uint32_t circle_log, uint32_t* tw_host, uint32_t* itw_host);

/* ------------------------------------------------------------------ field */
/* batch_inverse_base_field (cuda/ffi.zig:43); FieldOps.batchInverse (backend/field_ops.zig). */
int stwo_rocm_batch_inverse(stwo_rocm_ctx* ctx, const uint32_t* d_in, uint64_t n, uint32_t* d_out);
/* bit_reverse_base_field / bit_reverse_secure_field (cuda/ffi.zig:44-45): words_per_elem 1 or 4. */
int stwo_rocm_bit_reverse(stwo_rocm_ctx* ctx, uint32_t* d_vals, uint32_t log_size, uint32_t words_per_elem);
/* accumulate (cuda/ffi.zig:76; backend/accumulation_ops.zig): acc[i] += src[i] mod p. */
int stwo_rocm_accumulate(stwo_rocm_ctx* ctx, uint32_t* d_acc, const uint32_t* d_src, uint64_t n_words);

/* ------------------------------------------------------------------ circle polynomials */
/* interpolateCircleBuffers (prover/pcs/columns/circle_transforms.zig:248-250;
 * prover/poly/circle/transforms.zig:249-360); ntt_b2n (cuda/ffi.zig:52).
 * In place: evaluations on CanonicCoset(log_size).circleDomain() -> coefficients. */
int stwo_rocm_circle_interpolate(stwo_rocm_ctx* ctx, uint32_t* const* cols, uint32_t n_cols, uint32_t log_size,
                                 double* gpu_ms);
/* evaluateCircleBuffers (circle_transforms.zig:443-451; transforms.zig:132-246); ntt_n2b (cuda/ffi.zig:51).
 * coeffs (2^log_size each) -> evaluations on CanonicCoset(log_size + log_blowup), written to out. */
int stwo_rocm_circle_evaluate(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                              uint32_t log_blowup, uint32_t* const* out, double* gpu_ms);
/* batch_eval_at_points (cuda/ffi.zig:88); evaluateCoefficientPlans (prover/pcs/sampled_values.zig:67-80).
 * out_host[4*i..] = poly_i(point) for coefficient columns of 2^log_size. */
int stwo_rocm_eval_at_point(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, uint32_t n_cols, uint32_t log_size,
                            const uint32_t point[8], uint32_t* out_host);

/* ------------------------------------------------------------------ Merkle */
/* commitMerkle (backend/merkle_ops.zig:20-56; prover/vcs_lifted/prover.zig:58-68,276-340);
 * commit_on_first_layer_lifted (cuda/ffi.zig:272-278).  Lifted Blake2s Merkle with the
 * "leaf"/"node" 64-byte domain prefixes over columns of any power-of-two lengths >= 2.
 * The tree stays resident; root_out (32 bytes) may be NULL. */
int stwo_rocm_merkle_commit(stwo_rocm_ctx* ctx, const uint32_t* const* cols, const uint32_t* log_sizes,
                            uint32_t n_cols, stwo_rocm_tree** out, uint8_t* root_out, double* gpu_ms);
/* commit_on_layer_in_gpu (cuda/ffi.zig:70): one parent layer (2^log_size nodes) from 2^(log_size+1) children. */
int stwo_rocm_merkle_parent_layer(stwo_rocm_ctx* ctx, const uint8_t* d_children, uint32_t log_size,
                                  uint8_t* d_parents);
/* MerkleTree(H).root / maxLogSize / readHashes / deinit (merkle_ops.zig:27-40). */
int stwo_rocm_tree_root(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, uint8_t root_out[32]);
int stwo_rocm_tree_max_log(stwo_rocm_tree* tree, uint32_t* out);
int stwo_rocm_tree_read_hashes(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, uint32_t layer_log, const uint32_t* indices,
                               uint32_t n, uint8_t* out_host);
/* MerkleTree(H).decommit (prover/vcs_lifted/decommit.zig:56-155): hash witness for sorted
 * query positions. witness_out receives up to witness_cap hashes; *n_witness the count. */
int stwo_rocm_tree_decommit(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree, const uint64_t* positions, uint32_t n_positions,
                            uint8_t* witness_out, uint32_t witness_cap, uint32_t* n_witness);
int stwo_rocm_tree_free(stwo_rocm_ctx* ctx, stwo_rocm_tree* tree);

/* ------------------------------------------------------------------ FRI */
/* foldLine / foldLineN (backend/fri_ops.zig:8-17; core/fri/folding.zig:352-380); fold_line (cuda/ffi.zig:61).
 * d_src: 2^log_src QM31 (AoS) on LineDomain(Coset.halfOdds(log_src)); d_dst: 2^(log_src-1) QM31. */
int stwo_rocm_fri_fold_line(stwo_rocm_ctx* ctx, const uint32_t* d_src, uint32_t log_src, const uint32_t alpha[4],
                            uint32_t* d_dst);
/* foldCircleIntoLine (backend/fri_ops.zig; core/fri/folding.zig:558-602); fold_circle_into_line (cuda/ffi.zig:62).
 * src_coords: 4 coordinate columns of 2^log_src on CanonicCoset(log_src).circleDomain();
 * d_dst (in/out): 2^(log_src-1) QM31, dst = dst * alpha^2 + fold. */
int stwo_rocm_fri_fold_circle_into_line(stwo_rocm_ctx* ctx, const uint32_t* const* src_coords, uint32_t log_src,
                                        const uint32_t alpha[4], uint32_t* d_dst);

/* ------------------------------------------------------------------ quotients */
/* computeLazyQuotients / commitLazyMerkle (prover/fri.zig:341-363; prover/pcs/quotients/lazy_provider.zig:316-395);
 * accumulate_quotients (cuda/ffi.zig:94-108).
 * Columns (device, lengths 2^column_log_sizes[i]) are lifted to 2^lifting_log_size.
 * Samples are listed in tree/column/sample order: (column index, point (8 words), value (4 words)).
 * Random coefficient powers and point batches follow core/pcs/quotients/samples.zig:245-312.
 * out_coords: 4 device columns of 2^lifting_log_size.  When tree_out is non-NULL the 4
 * coordinate columns are also committed (fused leaf hashing) and the tree returned. */
int stwo_rocm_compute_quotients(stwo_rocm_ctx* ctx, uint32_t lifting_log_size, const uint32_t* const* columns,
                                const uint32_t* column_log_sizes, uint32_t n_columns, const uint32_t* sample_columns,
                                const uint32_t* sample_points, const uint32_t* sample_values, uint32_t n_samples,
                                const uint32_t random_coeff[4], uint32_t* const* out_coords,
                                stwo_rocm_tree** tree_out);

/* ------------------------------------------------------------------ resident hooks */
/* The optional hooks the generic prover probes with @hasDecl(B, ...) (src/prover/fri.zig:341-586,
 * src/prover/pcs/commit_dispatch.zig:34-90), in the shape of the Metal backend's
 * (src/backends/metal/commit_backend.zig:120-800).  Every hook leaves its results in HBM: buffers
 * the caller allocated with stwo_rocm_buffer_alloc (the ResidentStorage of
 * src/backend/resident_storage.zig:1-8) and resident trees.  A Zig RocmBackend wraps the device
 * pointers as SecureColumnByCoords / LineEvaluation with a ResidentStorage destroyFn that calls
 * stwo_rocm_buffer_free. */

/* Device-resident Blake2s Fiat-Shamir channel (Blake2sChannel, src/core/channel/blake2s.zig:33-134;
 * MerkleChannel.mixRoot, src/core/vcs_lifted/blake2_merkle.zig:237-249): digest and draw counter in
 * HBM, every mix and draw run on the device.  The FRI cascade hook advances it without a host
 * round trip, as the Metal cascade does with its channel_state words
 * (metal/commit_backend.zig:700-760). */
typedef struct stwo_rocm_channel stwo_rocm_channel;
/* digest NULL: the all-zero initial digest. */
int stwo_rocm_channel_create(stwo_rocm_ctx* ctx, const uint8_t digest[32], uint32_t n_draws,
                             stwo_rocm_channel** out);
int stwo_rocm_channel_destroy(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch);
/* digest and n_draws as the host channel would hold them (either may be NULL). */
int stwo_rocm_channel_state(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint8_t digest[32], uint32_t* n_draws);
/* mixU32s / mixU64 / mixFelts (blake2s.zig:52-88): digest <- Blake2s(digest || LE words), n_draws <- 0.
 * felts: n QM31 of 4 canonical words each. */
int stwo_rocm_channel_mix_u32s(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* words, uint64_t n);
int stwo_rocm_channel_mix_u64(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint64_t value);
int stwo_rocm_channel_mix_felts(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* felts, uint64_t n);
/* MerkleChannel.mixRoot (blake2_merkle.zig:237-249). */
int stwo_rocm_channel_mix_root(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint8_t root[32]);
/* drawSecureFelt (blake2s.zig:109-112 with drawBaseFelts' rejection of words >= 2p): 4 words. */
int stwo_rocm_channel_draw_secure_felt(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint32_t out[4]);
/* n_blocks calls of drawU32s (blake2s.zig:90-107): 8 words per block into out. */
int stwo_rocm_channel_draw_u32s(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, uint32_t n_blocks, uint32_t* out);

/* prepareAndCommitOwned (commit_dispatch.zig:34-71 -> MetalCommitBackend.prepareAndCommitOwned;
 * prover/pcs/columns/preparation.zig:113-210): evaluation columns (device, 2^log_sizes[i] words,
 * bit-reversed circle-domain order) are interpolated in place into their coefficients (retained,
 * scheme.zig:97 .always), extended to CanonicCoset(log + log_blowup) into ext_out[i] and
 * committed as one lifted Blake2s Merkle tree (columns of any sizes).  root_out may be NULL. */
int stwo_rocm_prepare_and_commit_owned(stwo_rocm_ctx* ctx, uint32_t* const* cols, const uint32_t* log_sizes,
                                       uint32_t n_cols, uint32_t log_blowup, uint32_t* const* ext_out,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out, double* gpu_ms);
/* prepareAndCommitPolys (commit_dispatch.zig:74-90; commit_polys.zig:10-67): the same from
 * coefficient columns (left unchanged). */
int stwo_rocm_prepare_and_commit_polys(stwo_rocm_ctx* ctx, const uint32_t* const* coeffs, const uint32_t* log_sizes,
                                       uint32_t n_cols, uint32_t log_blowup, uint32_t* const* ext_out,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out, double* gpu_ms);
/* secureColumnFromLine / secureColumnForMerkle (metal/commit_backend.zig:164-197; the Metal
 * qm31ToCoordinates kernel): n QM31 values (AoS) -> 4 coordinate columns (SoA). */
int stwo_rocm_secure_column_from_line(stwo_rocm_ctx* ctx, const uint32_t* d_line, uint64_t n,
                                      uint32_t* const* coords);
/* commitMerkle over the 4 coordinate columns of a line evaluation (fri.zig:497-517 with
 * secureColumnForMerkle), hashed straight from the AoS line: 2^log_size QM31. */
int stwo_rocm_commit_line(stwo_rocm_ctx* ctx, const uint32_t* d_line, uint32_t log_size, stwo_rocm_tree** tree_out,
                          uint8_t* root_out);
/* foldLineAndCommitNext (fri.zig:536-551; metal/commit_backend.zig:525-634): n_folds fold_line
 * steps of the line of log_src (alpha, then alpha^2, ...) into d_dst (2^(log_src - n_folds) QM31),
 * and the commitment of d_dst (its leaves hashed by the last fold). */
int stwo_rocm_fri_fold_line_and_commit(stwo_rocm_ctx* ctx, const uint32_t* d_src, uint32_t log_src,
                                       const uint32_t alpha[4], uint32_t n_folds, uint32_t* d_dst,
                                       stwo_rocm_tree** tree_out, uint8_t* root_out);
/* commitFriLayers / commitFriLineCascade (fri.zig:480-490; metal/commit_backend.zig:636-800),
 * fold_step 1: every inner layer from the line d_line (log_size) down to the last layer
 * (log_last_size >= 1), with the transcript on the device channel ch.  Per layer of log
 * l = log_size .. log_last_size+1: commit the line, mixRoot, draw the fold alpha, fold.
 * trees_out[k] (log_size - log_last_size handles) is the tree of the line of log log_size - k;
 * d_lines[k] (caller buffers, 2^(log_size-1-k) QM31) receives the line of log log_size-1-k for
 * k < log_size - log_last_size - 1; d_last_line (2^log_last_size QM31) the last layer's line. */
int stwo_rocm_fri_commit_layers(stwo_rocm_ctx* ctx, stwo_rocm_channel* ch, const uint32_t* d_line,
                                uint32_t log_size, uint32_t log_last_size, uint32_t* const* d_lines,
                                stwo_rocm_tree** trees_out, uint32_t* d_last_line);
/* Query-value gather of the decommitment (prover/vcs_lifted/decommit.zig:56-237; the Metal
 * decommit kernels, shaders/core/decommit.metal:11-334): out_host[(i * n_positions + q) *
 * words_per_value + w] = word w of column i at the lifted row of positions[q] (a tree of max log
 * lifting_log; lifted row ((p >> (lifting_log - log + 1)) << 1) + (p & 1), leaves.zig:105-111).
 * words_per_value 1: M31 columns; 4: QM31 lines (AoS). */
int stwo_rocm_gather_values(stwo_rocm_ctx* ctx, const uint32_t* const* cols, const uint32_t* log_sizes,
                            uint32_t n_cols, uint32_t words_per_value, uint32_t lifting_log,
                            const uint64_t* positions, uint32_t n_positions, uint32_t* out_host);

/* ------------------------------------------------------------------ wide Fibonacci */
/* fillQuadraticRecurrenceTrace (examples/wide_fibonacci/trace.zig:63-72); generate_wide_fibonacci_trace
 * (cuda/ffi.zig:263).  d_out: seq_len columns of 2^log_n_rows, column-major.
 * mode 0: seeds (1, logicalRow(i)); mode 1: seeded uniform M31 pair per row (seed). */
int stwo_rocm_widefib_trace(stwo_rocm_ctx* ctx, uint32_t log_n_rows, uint32_t seq_len, uint32_t mode, uint64_t seed,
                            uint32_t* d_out);
/* computeCompositionEvaluation (cpu_scalar/secure_composition.zig; wide_fibonacci/component.zig:146-228):
 * eval_cols = trace evaluations on CanonicCoset(log_n_rows + 1); out_coords = 4 device columns. */
int stwo_rocm_widefib_composition(stwo_rocm_ctx* ctx, const uint32_t* const* eval_cols, uint32_t n_cols,
                                  uint32_t log_n_rows, const uint32_t random_coeff[4], uint32_t* const* out_coords);

/* ------------------------------------------------------------------ proof of work */
/* Blake2sChannel.grind (core/channel/blake2s.zig:168-258): lowest nonce with >= pow_bits
 * trailing zeros of H(H(0x12345678 || 0^12 || digest || bits) || nonce). */
int stwo_rocm_pow_grind(stwo_rocm_ctx* ctx, const uint8_t digest[32], uint32_t pow_bits, uint64_t* nonce);

/* ------------------------------------------------------------------ whole proof */
typedef struct {
    uint32_t log_n_rows;
    uint32_t sequence_len;
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
    uint32_t trace_mode; /* 0: reference seeds (1, i); 1: seeded uniform pair */
    uint64_t trace_seed;
} stwo_rocm_widefib_params;

typedef struct {
    /* main_commit, composition, sampled_values, quotients, fri_commit, pow, decommit, encode */
    double stage_gpu_ms[8];
    double stage_host_ms[8];
    double total_gpu_ms;
    double total_host_ms;
    uint64_t peak_hbm_bytes;
} stwo_rocm_prove_stats;

/* ProverEngine.prove for the wide-Fibonacci example (examples/wide_fibonacci.zig:181-343 ->
 * examples/common/prover_transaction.zig:222-339 -> prover/prove.zig:209-389) with
 * PcsConfig{pow_bits, FriConfig{log_last_layer_degree_bound, log_blowup_factor, n_queries}}.
 * d_trace: optional resident trace (sequence_len columns of 2^log_n_rows, column-major);
 * NULL generates it on the device inside the call.  Writes the canonical proof bytes
 * (compact JSON wire, src/interop/proof_wire.zig:15-76) to proof_out; when cap is too
 * small returns STWO_ROCM_EINVAL with *proof_len set to the required size.  stats is optional:
 * NULL skips the per-stage timing events (each costs a few microseconds of GPU time). */
int stwo_rocm_prove_wide_fibonacci(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params,
                                   const uint32_t* d_trace, uint8_t* proof_out, uint64_t cap, uint64_t* proof_len,
                                   stwo_rocm_prove_stats* stats);

/* Many wide-Fibonacci proofs in single launches.  No single reference counterpart: the
 * reference proves one statement per ProverEngine.prove call (examples/wide_fibonacci.zig:181-343);
 * this entry point proves n_proofs statements of one shape and PCS configuration (params) at
 * once, every stage one launch over all of them, and each proof is byte-identical to what
 * stwo_rocm_prove_wide_fibonacci returns for the same parameters.  Proof k uses trace_seeds[k]
 * (trace_seeds NULL: params->trace_seed for all).  Limits: log_n_rows + log_blowup_factor <= 17,
 * log_last_layer_degree_bound + log_blowup_factor <= 10, n_proofs 1..65535, device memory
 * n_proofs * stats->slice_bytes.  The proofs are written back to back to proofs_out, proof k at
 * [offsets[k], offsets[k + 1]) (offsets has n_proofs + 1 entries); when cap is too small the call
 * returns STWO_ROCM_EINVAL with offsets[n_proofs] set to the bytes needed.  stats is optional. */
typedef struct {
    double device_ms;     /* first launch to the readback of every proof's FRI state */
    double tail_ms;       /* last layers, proof of work, decommitment gather and wire encoding */
    double total_ms;      /* the whole call */
    uint64_t slice_bytes; /* device bytes per proof */
    /* host wall time of the phases: set-up (tables, uploads, launches), wait for the device,
     * last layers, proof of work, query plans and gather, wire encoding */
    double phase_ms[6];
} stwo_rocm_batch_stats;
int stwo_rocm_prove_wide_fibonacci_batch(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params,
                                         uint32_t n_proofs, const uint64_t* trace_seeds, uint8_t* proofs_out,
                                         uint64_t cap, uint64_t* offsets, stwo_rocm_batch_stats* stats);
/* The same in two steps, for repeated batches of one shape (no reference counterpart; for
 * serving loops that prove many batches): create sets up the device allocation, tables and
 * pinned readback of n_proofs proofs of params' shape and PCS configuration; each run proves
 * one batch (trace_seeds as above; params->trace_seed when NULL) with the same output
 * conventions.  use_graph != 0 records the device part of the first run into a hipGraph and
 * replays it in later runs (runs eagerly when the context profiles kernels or the plan's
 * quotients need a temporary buffer: more than 120 columns on at most 2^16 rows).  A plan
 * belongs to its context; destroy it before the context. */
typedef struct stwo_rocm_batch_prover stwo_rocm_batch_prover;
int stwo_rocm_batch_prover_create(stwo_rocm_ctx* ctx, const stwo_rocm_widefib_params* params, uint32_t n_proofs,
                                  stwo_rocm_batch_prover** out);
int stwo_rocm_batch_prover_run(stwo_rocm_ctx* ctx, stwo_rocm_batch_prover* bp, const uint64_t* trace_seeds,
                               int use_graph, uint8_t* proofs_out, uint64_t cap, uint64_t* offsets,
                               stwo_rocm_batch_stats* stats);
int stwo_rocm_batch_prover_destroy(stwo_rocm_ctx* ctx, stwo_rocm_batch_prover* bp);

/* Poseidon AIR statement and PCS configuration (examples/poseidon/input.zig:20-22;
 * core/pcs/mod.zig PcsConfig). */
typedef struct {
    uint32_t log_n_instances; /* 2^log_n_instances permutations, 8 per trace row: 4..29 */
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
} stwo_rocm_poseidon_params;

/* poseidon.prove (examples/poseidon.zig:67-73 -> prover_transaction.zig -> prover/prove.zig):
 * the 1264-column Poseidon trace of input.zig:121-211 generated on the device, one main tree,
 * the constant composition of poseidon.zig:562-569, and the rest of the proof as for the
 * wide-Fibonacci AIR.  Same output, buffer and stats conventions. */
int stwo_rocm_prove_poseidon(stwo_rocm_ctx* ctx, const stwo_rocm_poseidon_params* params, uint8_t* proof_out,
                             uint64_t cap, uint64_t* proof_len, stwo_rocm_prove_stats* stats);

/* Periodic XOR example AIR statement and PCS configuration (examples/xor/input.zig:12-16). */
typedef struct {
    uint32_t log_size; /* 1..26 */
    uint32_t log_step; /* <= log_size */
    uint64_t offset;
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
} stwo_rocm_xor_params;

/* xor.prove (examples/xor.zig -> prover_transaction.zig -> prover/prove.zig): the IsFirst and
 * IsStepWithOffset indicators (preprocessed tree) and the main column of xor/input.zig:31-107
 * generated on the device, the constant composition of xor.zig:590-597, every column sampled
 * at the OODS point.  Same output, buffer and stats conventions. */
int stwo_rocm_prove_xor(stwo_rocm_ctx* ctx, const stwo_rocm_xor_params* params, uint8_t* proof_out, uint64_t cap,
                        uint64_t* proof_len, stwo_rocm_prove_stats* stats);

/* XOR-lookup AIR statement and PCS configuration (the AIR of this backend's LogUp benchmark; see
 * stwo_rocm_prove_logup_xor). */
typedef struct {
    uint32_t log_n; /* 16..26 */
    uint32_t pow_bits;
    uint64_t seed;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
    uint32_t reserved;
} stwo_rocm_logup_params;

/* No single reference counterpart: the XOR-lookup AIR of the LogUp benchmark as a whole proof
 * through the generic transaction (prover_transaction.zig:222-339 -> prove.zig:209-389 ->
 * scheme.zig:571-800), with the LogUp machinery of src/frontends/riscv/air/logup.zig:74-105: tree 0
 * is_first and the table tuple (x, y, x ^ y); tree 1 the lookups a, b (uniform 8-bit from the
 * seed), c = a ^ b and the table multiplicities; tree 2 the pair-batched cumulative column of
 * 1 / (z - (a + alpha b + alpha^2 c)) - m / (z - (x + alpha y + alpha^2 (x ^ y))), sampled at the
 * OODS point and at the previous row; tree 3 the composition of the pair constraint.  The claimed
 * sum is zero (the lookups balance the table).  Same output and buffer conventions as
 * stwo_rocm_prove_wide_fibonacci (no stage timings). */
int stwo_rocm_prove_logup_xor(stwo_rocm_ctx* ctx, const stwo_rocm_logup_params* params, uint8_t* proof_out,
                              uint64_t cap, uint64_t* proof_len);

/* ------------------------------------------------------------------ LogUp / GKR */
/* Layer kinds of src/prover/lookups/gkr_circuit.zig:79-92. */
#define STWO_ROCM_GKR_GRAND_PRODUCT 0
#define STWO_ROCM_GKR_LOGUP_GENERIC 1       /* numerators QM31 (AoS, 4 words) */
#define STWO_ROCM_GKR_LOGUP_MULTIPLICITIES 2 /* numerators M31 (1 word) */
#define STWO_ROCM_GKR_LOGUP_SINGLES 3       /* numerators all one (d_num ignored) */

/* XOR-lookup AIR of the LogUp workload (no reference counterpart: the AIR is defined by this
 * backend's LogUp benchmark): 2^log_size rows of a, b uniform 8-bit (splitmix64 of
 * seed ^ i*0xD1B54A32D192ED03), c = a ^ b, and mult[x*256+y] = occurrences of (x, y). */
int stwo_rocm_xor_trace(stwo_rocm_ctx* ctx, uint32_t log_size, uint64_t seed, uint32_t* d_a, uint32_t* d_b,
                        uint32_t* d_c, uint32_t* d_mult);
/* den[i] = z - (a + b*alpha + c*alpha^2) as QM31 AoS (the relation combination of
 * src/frontends/riscv/air/logup.zig:131-175 for a 3-tuple); with d_a = d_b = d_c = NULL the rows are
 * the table ((i >> 8) & 255, i & 255, their xor), repeating every 2^16 rows. */
int stwo_rocm_logup_denominators(stwo_rocm_ctx* ctx, const uint32_t* d_a, const uint32_t* d_b, const uint32_t* d_c,
                                 uint64_t n, const uint32_t z[4], const uint32_t alpha[4], uint32_t* d_den);
/* Cumulative LogUp column in trace order, replaces cumulativeColumn
 * (src/frontends/riscv/air/logup.zig:74-91): d_sums[i] = sum_{j<=i} num_j / den_j (QM31 AoS),
 * claimed = the total.  num_kind: 0 QM31 AoS, 1 M31, 2 all ones (d_num ignored).  A zero
 * denominator returns STWO_ROCM_EINVAL. */
int stwo_rocm_logup_cumulative(stwo_rocm_ctx* ctx, int num_kind, const uint32_t* d_num, const uint32_t* d_den,
                               uint64_t n, uint32_t* d_sums, uint32_t claimed[4]);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same result as
 * stwo_rocm_logup_cumulative, with no allocation: the scan's block partial sums, the total and
 * the zero-denominator flag live in a fixed scratch array of the library on each device
 * (1 MiB + 256 B, held until the process ends).  n <= 2^27 rows, else STWO_ROCM_EINVAL.  Every
 * call on a device shares that device's array: calls that overlap on one GPU (from several
 * contexts or threads) race. */
int stwo_rocm_logup_cumulative_noalloc(stwo_rocm_ctx* ctx, int num_kind, const uint32_t* d_num,
                                       const uint32_t* d_den, uint64_t n, uint32_t* d_sums, uint32_t claimed[4]);
/* GkrOps.genEqEvals (src/backend/gkr_ops.zig:1-23; gkr_circuit.zig:651-674): d_out[i] (2^n QM31) =
 * scale * prod_j (bit_{n-1-j}(i) ? y_j : 1 - y_j); y is a host array of n QM31. */
int stwo_rocm_gkr_gen_eq_evals(stwo_rocm_ctx* ctx, const uint32_t* y, uint32_t n, const uint32_t scale[4],
                               uint32_t* d_out);
/* GkrOps.nextLayer (gkr_circuit.zig:139-160,449-503): halves a layer of n_vars variables
 * (pairs 2i, 2i+1); LogUp layers come out generic (QM31 numerators). */
int stwo_rocm_gkr_next_layer(stwo_rocm_ctx* ctx, int kind, const uint32_t* d_num, const uint32_t* d_den,
                             uint32_t n_vars, uint32_t* d_out_num, uint32_t* d_out_den);
/* GkrOps.sumAsPolyInFirstVariable (gkr_circuit.zig:261-316,390-429) of the sum-check oracle over a layer of
 * n_vars variables with eq table d_eq, batching lambda, claim, eq point y (host, y_len QM31) and
 * eq_fixed_var_correction: up to 4 QM31 coefficients (ascending, trailing zeros trimmed). */
int stwo_rocm_gkr_sum_as_poly_in_first_variable(stwo_rocm_ctx* ctx, int kind, const uint32_t* d_eq,
                                                const uint32_t* d_num, const uint32_t* d_den, uint32_t n_vars,
                                                const uint32_t lambda[4], const uint32_t claim[4], const uint32_t* y,
                                                uint32_t y_len, const uint32_t correction[4], uint32_t coeffs_out[16],
                                                uint32_t* n_coeffs);
/* Mle.fixFirstVariable (src/prover/lookups/mle.zig:80-95): d_out[i] = r*(in[i+h]-in[i]) + in[i]
 * (QM31 AoS out; input M31 when base_field != 0). */
int stwo_rocm_mle_fix_first_variable(stwo_rocm_ctx* ctx, int base_field, const uint32_t* d_in, uint32_t n_vars,
                                     const uint32_t r[4], uint32_t* d_out);
/* Batch GKR proof, replaces proveBatch (src/prover/lookups/gkr_prover.zig:34-298 with
 * sumcheck.zig:76-197) on a Blake2s channel starting at channel_digest.  Writes the proof word
 * stream: per GKR layer [n_rounds, each (len, len QM31)], [n_masks, each (instance, n_cols,
 * 2*n_cols QM31)]; per instance [n, output claims]; [n, ood point]; per instance [n, claims to
 * verify], n_vars; then the final channel digest (8 words).  When cap is too small returns
 * STWO_ROCM_EINVAL with *n_words set. */
int stwo_rocm_gkr_prove_batch(stwo_rocm_ctx* ctx, uint32_t n_instances, const int* kinds,
                              const uint32_t* const* d_nums, const uint32_t* const* d_dens, const uint32_t* n_vars,
                              const uint8_t channel_digest[32], uint32_t* words_out, uint64_t cap,
                              uint64_t* n_words);

/* LogUp workload end to end (no single reference counterpart: the sequence of commit
 * (src/prover/pcs/commit_polys.zig:10-67), LogUp cumulative columns (logup.zig:74-91), batch GKR
 * (gkr_prover.zig:34-298), sampled values (sampled_values.zig:67-80) and quotient accumulation
 * (row_evaluation.zig:402-521) applied to the XOR-lookup AIR of stwo_rocm_xor_trace on 2^log_n rows,
 * log_n in 16..26).  Writes: main, interaction and quotient Merkle roots (3 x 8 words), claimed sums
 * of lookups and table (2 x 4), 12 sampled values (12 x 4), the GKR word count and proof words, and
 * the final channel digest (8).  total_ms (optional): host wall time of the call.  Returns
 * STWO_ROCM_EINVAL if the lookups do not balance or cap is too small (*n_words set). */
int stwo_rocm_logup_xor_pipeline(stwo_rocm_ctx* ctx, uint32_t log_n, uint64_t seed, uint32_t* words_out,
                                 uint64_t cap, uint64_t* n_words, double* total_ms);

/* ------------------------------------------------------------------ batched pipeline */
/* No single reference counterpart: the batching pattern around the reference's per-column
 * hooks interpolateCircleBuffers / evaluateCircleBuffers (src/prover/pcs/columns/
 * circle_transforms.zig:248-250,443-451), commitMerkle (src/backend/merkle_ops.zig:20-56) and
 * foldLine (src/prover/fri.zig:434-586).  n_columns seeded uniform-M31 columns with log sizes drawn
 * from [log_min, log_max] are interpolated, LDE'd at blowup 2^log_blowup and committed in one
 * lifted Merkle tree; then n_folds FRI fold_line rounds run on a seeded QM31 line evaluation of log
 * fri_log, each layer committed and its root mixed into a device Fiat-Shamir channel that draws the
 * fold alpha.  All buffers and descriptor tables are planned at create time and the launch
 * sequence is captured as one hipGraph.  Result words: main root, n_folds + 1 line roots (8 words
 * each), final channel digest (8), draw counter. */
typedef struct stwo_rocm_batched stwo_rocm_batched;
typedef struct {
    uint32_t n_columns, log_min, log_max, log_blowup, fri_log, n_folds;
    uint64_t seed;
} stwo_rocm_batched_spec;
int stwo_rocm_batched_create(stwo_rocm_ctx* ctx, const stwo_rocm_batched_spec* spec, stwo_rocm_batched** out);
/* Launch-shape tuning (results are bit-identical for every setting): FFT LDS tile 2^11..2^13 words,
 * Merkle parent levels per launch 2..10.  Applies to this pipeline only (the context's other work
 * keeps its shapes) and recaptures the graph.  A replay recaptures by itself when the context's
 * twiddle tower has moved since the capture (a larger transform on the context, twiddles_prepare,
 * release_cached). */
int stwo_rocm_batched_set_tuning(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t fft_tile_log,
                                 uint32_t merkle_levels);
/* Times every (fft_tile_log in 11..13) x (merkle_levels in {6, 8, 10}) graph over reps replays, checks
 * that all candidates produce the same result words, keeps the fastest; json_out (optional) gets
 * [{fft_tile_log, merkle_levels, ms}, ...]. */
int stwo_rocm_batched_autotune(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t reps, uint32_t* fft_tile_log,
                               uint32_t* merkle_levels, char* json_out, size_t cap);
/* n_runs replays of the graph (use_graph != 0) or of the same launches issued eagerly; device ms per
 * run; the result words of the last run. */
int stwo_rocm_batched_run(stwo_rocm_ctx* ctx, stwo_rocm_batched* b, uint32_t n_runs, int use_graph,
                          double* ms_per_run, uint32_t* words_out, uint64_t cap, uint64_t* n_words);
int stwo_rocm_batched_destroy(stwo_rocm_ctx* ctx, stwo_rocm_batched* b);

/* ------------------------------------------------------------------ serving */
/* Continuous-batching proof service, extending the reference's one-job-at-a-time resource
 * admission (src/prover/native/resource_admission.zig) and its prover session loop
 * (src/tools/metal_prover_session/app.zig): n_slots proof slots (1..16), each with its own device
 * context and HIP stream; requests become eligible at arrival_ms (relative to the call, sorted) and
 * the FIFO head is admitted when a slot is idle and it fits hbm_budget_bytes (0 = unlimited): the
 * device memory every slot holds while idle (its tower, tables and cached arena blocks), plus the
 * footprint estimates of the jobs in flight and of the head job, must not exceed the budget; an
 * idle slot trims its cached blocks, and running slots trim theirs when their jobs end, until the
 * head fits.  stats->peak_device_bytes reports the device memory the slots really held.  Jobs on
 * different slots overlap on the device.  Job kinds: 0
 * wide-Fibonacci proof (functional protocol, seeded uniform trace: trace_mode 1 of
 * stwo_rocm_prove_wide_fibonacci), 1 XOR-lookup LogUp workload (stwo_rocm_logup_xor_pipeline,
 * log_n 16..24), 2 Poseidon AIR proof (stwo_rocm_prove_poseidon, log_n = log_n_instances 4..24,
 * functional protocol), 3 periodic XOR AIR proof (stwo_rocm_prove_xor, log_n = log_size 2..24,
 * seq_len = log_step, seed = offset, functional protocol), 4 XOR-lookup AIR proof
 * (stwo_rocm_prove_logup_xor, log_n 16..24, functional protocol).  Each result carries the
 * Blake2s-256 digest of its output bytes. */
typedef struct {
    uint32_t kind, log_n, seq_len, reserved;
    uint64_t seed;
    double arrival_ms;
} stwo_rocm_serve_job;
typedef struct {
    uint8_t digest[32];
    double arrival_ms, start_ms, end_ms;
    uint32_t slot, status;
    uint64_t output_bytes;
} stwo_rocm_serve_result;
typedef struct {
    double makespan_ms, p50_ms, p99_ms, mean_ms, jobs_per_s;
    uint64_t peak_admitted_bytes;
    uint64_t peak_device_bytes; /* device memory held by all slots (live + cached arena blocks), peak */
    uint32_t max_in_flight, n_slots;
} stwo_rocm_serve_stats;
int stwo_rocm_serve(stwo_rocm_ctx* ctx, const stwo_rocm_serve_job* jobs, uint32_t n_jobs, uint32_t n_slots,
                    uint64_t hbm_budget_bytes, stwo_rocm_serve_result* results, stwo_rocm_serve_stats* stats);
/* The device-byte footprint estimate that stwo_rocm_serve admits a job against (kind, log_n and
 * seq_len as above; the other fields are ignored).  It takes the place of the reference's
 * host-byte accounting per committed cell (src/prover/native/resource_admission.zig:5-109) for a
 * device backend; it needs no device and no context. */
int stwo_rocm_job_device_bytes(const stwo_rocm_serve_job* job, uint64_t* bytes);

/* ------------------------------------------------------------------ measurement */
/* No counterpart in the reference: a measurement hook for callers that need per-kernel
 * device time (the reference's stage profiler, src/prover/stage_profile.zig, times stages
 * on the host).  While enabled, every kernel launch of the context is bracketed by HIP
 * events on the context stream.  profile_read synchronises, returns a JSON object
 * {class: {launches, ms, max_ms, bytes}} (bytes = algorithmic HBM bytes of the launches)
 * and clears the records.  profile_select restricts the records to the kernel classes named
 * in kernel_class, separated by '|' (NULL or "" = every class), so a timed run can carry
 * events around its dominant kernels only; profile_sample records only every k-th matching launch (k >= 1, default 1),
 * for a class launched many times per run. */
int stwo_rocm_profile_enable(stwo_rocm_ctx* ctx, int enable);
int stwo_rocm_profile_select(stwo_rocm_ctx* ctx, const char* kernel_class);
int stwo_rocm_profile_sample(stwo_rocm_ctx* ctx, uint32_t every);
int stwo_rocm_profile_read(stwo_rocm_ctx* ctx, char* json_out, size_t cap, size_t* len);

#ifdef __cplusplus
}
#endif
#endif
