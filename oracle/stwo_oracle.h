/*
 * stwo_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of the reference (welttowelt/stwo-zig) CPU backend for the
 * wide-Fibonacci proving path.  It is the checker for the HIP backend: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * The product library (stwo-zig_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  All arithmetic is
 * exact (M31 / CM31 / QM31), so results must be bit-identical to the reference.
 */
#ifndef STWO_ORACLE_H
#define STWO_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORA_P 0x7fffffffu

/* ---- fields (src/core/fields/{m31,cm31,qm31}.zig) ---- */
uint32_t ora_m31_mul(uint32_t a, uint32_t b);
uint32_t ora_m31_inv(uint32_t a);
void ora_qm31_mul(const uint32_t a[4], const uint32_t b[4], uint32_t out[4]);
void ora_qm31_inv(const uint32_t a[4], uint32_t out[4]);

/* ---- circle group (src/core/circle.zig) ---- */
/* Point of the M31 circle group at index `idx` (mod 2^31): out = (x, y). */
void ora_circle_point_at_index(uint32_t idx, uint32_t out[2]);
/* Point i of CanonicCoset(log_size).circleDomain() (src/core/poly/circle/domain.zig:58-62). */
void ora_circle_domain_point(uint32_t log_size, uint64_t i, uint32_t out[2]);

/* ---- twiddles (src/prover/poly/twiddles.zig:37-91) ---- */
/* Writes 2^(log-1) twiddles and inverse twiddles of CanonicCoset(log). */
void ora_twiddles(uint32_t circle_log, uint32_t *tw, uint32_t *itw);

/* ---- circle FFT (src/prover/poly/circle/transforms.zig:149-360) ---- */
void ora_circle_evaluate(uint32_t *values, uint32_t log_size);
void ora_circle_interpolate(uint32_t *values, uint32_t log_size);
/* coeffs (2^log) -> evaluations on CanonicCoset(log + blowup) (zero padded). */
void ora_circle_lde(const uint32_t *coeffs, uint32_t log_size, uint32_t log_blowup, uint32_t *out);

/* ---- point evaluation (src/prover/poly/circle/point_evaluation.zig:14-38,469-488) ---- */
void ora_eval_at_point(const uint32_t *coeffs, uint32_t log_size, const uint32_t point[8], uint32_t out[4]);

/* ---- Blake2s (src/core/crypto/blake2s_backend.zig; RFC 7693) ---- */
void ora_blake2s(const uint8_t *data, size_t len, uint8_t out[32]);

/* ---- lifted Merkle (src/prover/vcs_lifted/{prover,leaves,layers}.zig) ---- */
typedef struct ora_tree ora_tree;
ora_tree *ora_merkle_commit(const uint32_t *const *cols, const uint32_t *log_sizes, uint32_t n_cols);
void ora_tree_root(const ora_tree *t, uint8_t out[32]);
uint32_t ora_tree_max_log(const ora_tree *t);
/* Copies layer `layer_log` (2^layer_log hashes x 32 bytes) into out. */
void ora_tree_layer(const ora_tree *t, uint32_t layer_log, uint8_t *out);
void ora_tree_free(ora_tree *t);

/* ---- FRI folds (src/core/fri/folding.zig:352-380,558-602) ---- */
/* src: 2^log_src QM31 (AoS, 4 words each) on LineDomain(coset(initial_index, log_src));
 * dst: 2^(log_src-1) QM31. */
void ora_fold_line(const uint32_t *src, uint32_t log_src, uint32_t coset_initial_index,
                   const uint32_t alpha[4], uint32_t *dst);
/* src: 4 coordinate columns of 2^log_src on CanonicCoset(log_src).circleDomain();
 * dst (in/out): 2^(log_src-1) QM31, dst = dst*alpha^2 + fold. */
void ora_fold_circle_into_line(const uint32_t *const src[4], uint32_t log_src, const uint32_t alpha[4],
                               uint32_t *dst);

/* ---- channel PoW (src/core/channel/blake2s.zig:135-258) ---- */
uint64_t ora_pow_grind(const uint8_t digest[32], uint32_t pow_bits);

/* ---- wide Fibonacci ---- */
/* Trace in bit-reversed circle-domain order (src/examples/wide_fibonacci/trace.zig:40-162).
 * mode 0: reference seeds (1, logicalRow(i)); mode 1: seeded uniform M31 initial pair. */
void ora_widefib_trace(uint32_t log_n_rows, uint32_t seq_len, uint32_t mode, uint64_t seed, uint32_t *out);

typedef struct {
    uint32_t log_n_rows;
    uint32_t sequence_len;
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
    uint32_t trace_mode;
    uint64_t trace_seed;
} ora_widefib_params;

/* Full proof: returns 0 on success and writes the canonical proof bytes (compact
 * JSON wire, src/interop/proof_wire.zig:15-76).  *out_len receives the length;
 * when cap is too small returns 2 and only sets *out_len. */
int ora_prove_wide_fibonacci(const ora_widefib_params *p, uint8_t *out, size_t cap, size_t *out_len);

/* Poseidon AIR trace (src/examples/poseidon/input.zig:121-158): 1264 columns of 2^log_n_rows. */
void ora_poseidon_trace(uint32_t log_n_rows, uint32_t *out);

/* Poseidon AIR proof (src/examples/poseidon.zig:67-73) for 2^log_n_instances permutations
 * (8 per row); prm supplies the PCS configuration only.  Same return convention. */
int ora_prove_poseidon(uint32_t log_n_instances, const ora_widefib_params *prm, uint8_t *out, size_t cap,
                       size_t *out_len);

/* Periodic XOR example AIR columns (src/examples/xor/input.zig:31-107), 2^log_size words each. */
void ora_xor_example_columns(uint32_t log_size, uint32_t log_step, uint64_t offset, uint32_t *is_first,
                             uint32_t *is_step, uint32_t *main_col);

/* Periodic XOR example AIR proof (src/examples/xor.zig); prm supplies the PCS configuration only. */
int ora_prove_xor(uint32_t log_size, uint32_t log_step, uint64_t offset, const ora_widefib_params *prm,
                  uint8_t *out, size_t cap, size_t *out_len);

/* Stage timings of the last ora_prove_wide_fibonacci call (milliseconds). */
void ora_last_timings(double out[8]);
void ora_set_threads(int n);
int ora_get_threads(void);

/* LogUp / GKR (see stwo_oracle.c). */
void ora_xor_trace(uint32_t log_size, uint64_t seed, uint32_t *a, uint32_t *b, uint32_t *c, uint32_t *mult);
void ora_logup_denominators(const uint32_t *a, const uint32_t *b, const uint32_t *c, uint64_t n, const uint32_t z[4],
                            const uint32_t alpha[4], uint32_t *den);
int ora_logup_cumulative(int num_kind, const uint32_t *num, const uint32_t *den, uint64_t n, uint32_t *sums,
                         uint32_t claimed[4]);
void ora_gkr_gen_eq_evals(const uint32_t *y, uint32_t n, const uint32_t scale[4], uint32_t *out);
void ora_gkr_next_layer(int kind, const uint32_t *num, const uint32_t *den, uint32_t n_vars, uint32_t *out_num,
                        uint32_t *out_den);
void ora_mle_fix_first(int base, const uint32_t *in, uint32_t n_vars, const uint32_t assignment[4], uint32_t *out);
size_t ora_gkr_prove(uint32_t n_inst, const int *kinds, const uint32_t *const *nums, const uint32_t *const *dens,
                     const uint32_t *n_vars, const uint8_t digest_in[32], uint32_t *out, size_t cap);
size_t ora_logup_xor_pipeline(uint32_t log_n, uint64_t seed, uint32_t *out, size_t cap);
size_t ora_batched_pipeline(uint32_t n_columns, uint32_t log_min, uint32_t log_max, uint32_t log_blowup,
                            uint32_t fri_log, uint32_t n_folds, uint64_t seed, uint32_t *out, size_t cap);
/* XOR-lookup AIR as a whole proof (tree 0 is_first + table tuple, tree 1 a, b, c, multiplicities,
 * tree 2 the pair-batched cumulative column, tree 3 the composition of the pair constraint);
 * prm supplies the PCS configuration.  Returns 5 if the lookups do not balance. */
int ora_prove_logup_xor(uint32_t log_n, uint64_t seed, const ora_widefib_params *prm, uint8_t *out, size_t cap,
                        size_t *out_len);


#ifdef __cplusplus
}
#endif
#endif
