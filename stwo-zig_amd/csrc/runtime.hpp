// Internal runtime of the rocm backend: device context, caching device arena,
// resident twiddle tower, resident Merkle trees and the launch wrappers of
// every kernel.  Not part of the C-ABI (include/stwo_rocm.h is).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "blake2s.hpp"
#include "fields.hpp"

namespace sr {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define SR_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t _e = (call);                                                                        \
        if (_e != hipSuccess)                                                                          \
            throw ::sr::Error(3, std::string(#call " failed: ") + hipGetErrorString(_e) + " at " +     \
                                     __FILE__ + ":" + std::to_string(__LINE__));                       \
    } while (0)

#define SR_CHECK(cond, msg)                                \
    do {                                                   \
        if (!(cond)) throw ::sr::Error(1, std::string(msg)); \
    } while (0)

// Lifted Merkle tree resident in HBM: layer k (2^k hashes) at hash offset
// 2^k - 1 (root first, as MerkleProverLifted.layers, prover.zig:23).
struct DeviceTree {
    uint32_t max_log = 0;
    uint8_t* layers = nullptr;  // device, (2^(max_log+1)-1) * 32 bytes
    uint8_t root[32];
    uint8_t* layer(uint32_t k) const { return layers + (((uint64_t)1 << k) - 1) * 32; }
};

// Fiat-Shamir channel state kept in HBM during the FRI commit loop.
struct DevChannel {
    uint32_t digest[8];
    uint32_t n_draws;
    uint32_t status;     // non-zero: draw rejection cap reached
    uint32_t alpha[4];   // last drawn secure felt
    uint32_t alpha2[4];  // its square
};

// Device bytes held (live + cached arena blocks) by a set of contexts that share one HBM
// budget (the serving slots), with the peak since the last reset.
struct MemLedger {
    std::atomic<uint64_t> held{0}, peak{0};
    void add(uint64_t b) {
        const uint64_t h = held.fetch_add(b) + b;
        uint64_t p = peak.load();
        while (h > p && !peak.compare_exchange_weak(p, h)) {
        }
    }
    void sub(uint64_t b) { held.fetch_sub(b); }
    void reset_peak() { peak.store(held.load()); }
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string last_error;
    uint64_t hbm_budget = 0;  // 0 = unlimited
    uint64_t bytes_live = 0, bytes_cached = 0, peak_live = 0;
    std::multimap<size_t, void*> free_blocks;
    std::unordered_map<void*, size_t> live_blocks;
    MemLedger* ledger = nullptr;  // shared count of device bytes held (set before ctx_init), or none

    // Twiddle tower of CanonicCoset(tower_log) (twiddle_tower.zig:29-115).
    uint32_t tower_log = 0;
    uint64_t tower_builds = 0;  // builds of the tower since ctx_init
    uint32_t* d_tw = nullptr;
    uint32_t* d_itw = nullptr;
    // The same towers doubled (2t as a plain u32 word, < 2^32): the register FFT multiplies by
    // 2t so that b*t mod p reduces with one shift (mmul_t2).
    uint32_t* d_tw2 = nullptr;
    uint32_t* d_itw2 = nullptr;
    // Completion counters of the tree-layer launches (zero between launches).
    uint32_t* d_tree_cnt = nullptr;
    // Domain tables per circle log: bit-reversed domain points and inverse y of the half coset.
    std::map<uint32_t, uint32_t*> domain_xy;     // 2 * 2^log words: x[] then y[]
    std::map<uint32_t, uint32_t*> circle_inv_y;  // 2^(log-1) words
    B2State leaf_seed, node_seed;
    GenDoubles gen;

    // pinned host staging: a readback buffer and a ring for small uploads.
    uint8_t* h_pinned = nullptr;
    size_t h_pinned_size = 0;
    uint8_t* h_ring = nullptr;
    size_t ring_cap = 0, ring_off = 0;

    // events for GPU timing
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_sync = nullptr;  // sync(): recorded, then polled (no blocking wait)

    // Opt-in per-kernel profile: events around each launch on `stream`,
    // aggregated by kernel class with the launch's algorithmic HBM bytes.
    struct KRec {
        std::string cls;
        hipEvent_t a, b;
        double bytes;
    };
    // Tuning parameters (results are bit-identical for every setting):
    uint32_t fft_tile_log = 13;   // FFT LDS tile, 2^11..2^13 words
    uint32_t merkle_levels = 10;  // Merkle parent levels per launch, 2..10
    bool profiling = false;
    std::vector<std::string> profile_only;  // non-empty: record only launches of these classes
    uint32_t profile_every = 1;  // record every k-th launch that matches
    uint64_t profile_seen = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_used = 0;
    std::vector<KRec> krecs;
    hipEvent_t pool_event();

    void* alloc(size_t bytes);
    void free(void* p);
    void trim();
    void sync();
    uint8_t* pinned(size_t bytes);
    // Copies `bytes` into the pinned upload ring and returns the pinned copy; the
    // slot stays valid until the next sync (the ring syncs the stream when it wraps).
    void* stage(const void* src, size_t bytes);
    void ensure_tower(uint32_t log);
    const uint32_t* twiddle_layer(bool inverse, uint32_t m, uint32_t s) const;
    const uint32_t* domain_points(uint32_t log);
    const uint32_t* inv_circle_y(uint32_t log);
};

// Context set-up on `device` (stream at default priority, hash seeds, circle generator) and
// release of every resource it holds (device blocks, pinned staging, events, stream).
void ctx_init(Ctx& c, int device, uint64_t hbm_budget);
void ctx_release(Ctx& c);

// Records one kernel launch in the context profile (no-op unless profiling).
struct KTimer {
    Ctx& c;
    std::string cls;
    double bytes;
    hipEvent_t a = nullptr;
    KTimer(Ctx& cc, std::string k, double b) : c(cc), cls(std::move(k)), bytes(b) {
        if (c.profiling && (c.profile_only.empty() || selected()) &&
            c.profile_seen++ % c.profile_every == 0) {
            a = c.pool_event();
            (void)hipEventRecord(a, c.stream);
        }
    }
    bool selected() const {
        for (const auto& k : c.profile_only)
            if (k == cls) return true;
        return false;
    }
    ~KTimer() {
        if (a) {
            hipEvent_t e = c.pool_event();
            (void)hipEventRecord(e, c.stream);
            c.krecs.push_back({cls, a, e, bytes});
        }
    }
};

// RAII device buffer from the context arena.
struct DBuf {
    Ctx* ctx = nullptr;
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(Ctx* c, size_t b) : ctx(c), p(b ? c->alloc(b) : nullptr), bytes(b) {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : ctx(o.ctx), p(o.p), bytes(o.bytes) { o.p = nullptr; }
    DBuf& operator=(DBuf&& o) noexcept {
        reset();
        ctx = o.ctx;
        p = o.p;
        bytes = o.bytes;
        o.p = nullptr;
        return *this;
    }
    ~DBuf() { reset(); }
    void reset() {
        if (p && ctx) ctx->free(p);
        p = nullptr;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

// ---- launch wrappers (kernels_*.hip) ----
// Circle FFT over columns given by device pointer arrays.  Forward evaluates
// coefficients (length 2^src_log, zero padded) on CanonicCoset(m); inverse
// interpolates evaluations of CanonicCoset(m) (src_log == m) and scales by 2^-m.
void launch_circle_fft(Ctx& c, bool inverse, const uint32_t* const* d_src, uint32_t* const* d_dst, uint32_t n_cols,
                       uint32_t m, uint32_t src_log);
// Interpolation of evaluations on CanonicCoset(n) into d_coef (the retained coefficients) and
// their extension to CanonicCoset(n + 1) into d_lde (device pointer arrays, one entry per
// column): the top strided passes of the two transforms run as one launch when the pass plans
// allow it, otherwise the two transforms run one after the other.  Same results either way.
void launch_circle_ifft_lde(Ctx& c, const uint32_t* const* d_eval, uint32_t* const* d_coef, uint32_t* const* d_lde,
                            uint32_t n_cols, uint32_t n);
// Lifted Merkle: columns sorted by log size ascending (d_cols/d_logs device arrays).
// Column descriptor of the leaf kernel; tables are padded to a multiple of 16 entries with
// mask 0 (padding reads column 0 and contributes a zero word).
struct LeafColDev {
    const uint32_t* col;
    uint32_t shift;  // lift shift level - log + 1 for the level of the column's segment (1 = full size)
    uint32_t mask;   // 0xffffffff, or 0 for padding
};
static_assert(sizeof(LeafColDev) == 16, "leaf column layout");
// Message blocks [blk0, blk0 + nblk) of the leaf hash, absorbed over 2^level lifted rows.
struct LeafSeg {
    uint32_t blk0, nblk, level;
    uint64_t col_words;  // sum of the segment's column sizes (words): algorithmic byte count only
    bool ident = true;   // every real column of the segment has the segment's log (row p reads row p)
};
// Descriptor table (16 entries per message block) and the segment plan for columns sorted
// by log size; every column's shift is relative to its segment's level.
struct LeafPlan {
    std::vector<LeafColDev> tab;
    std::vector<LeafSeg> segs;
};
LeafPlan merkle_leaf_plan(const uint32_t* const* h_cols, const uint32_t* h_logs, uint32_t n_cols);
// Leaves of t from a device-resident copy of plan.tab (no host traffic: capturable in a graph).
// Intermediate segment states live in t's two parent layers below the leaves until the
// parents are built.
void launch_merkle_leaves_dev(Ctx& c, const LeafColDev* d_tab, const std::vector<LeafSeg>& segs, uint32_t n_cols,
                              DeviceTree& t, uint32_t batch = 1, uint64_t bstride = 0);
// Host arrays in: column device pointers and log sizes, sorted by log size; max log = t.max_log.
void launch_merkle_leaves(Ctx& c, const uint32_t* const* h_cols, const uint32_t* h_logs, uint32_t n_cols,
                          DeviceTree& t);
void launch_merkle_leaves_qm31(Ctx& c, const uint32_t* d_qm31, uint32_t log_size, uint8_t* d_leaves);
// All parent layers above t's leaves.  With d_ch, the launch reaching the root
// also mixes it into the device channel and draws the next alpha (root copied to d_root_out).
// Batches (batch > 1) bring their own completion counters: d_cnt, TREE_CNT_WORDS per proof, zeroed.
void launch_merkle_parents(Ctx& c, DeviceTree& t, DevChannel* d_ch = nullptr, uint8_t* d_root_out = nullptr,
                           uint32_t batch = 1, uint64_t bstride = 0, uint32_t* d_cnt = nullptr);
void launch_merkle_parent_layer(Ctx& c, const uint8_t* d_children, uint32_t log_size, uint8_t* d_parents);
// FRI folds (folding.zig): line domain halfOdds(log_src) / circle domain CanonicCoset(log_src).
// d_alpha (optional): read alpha (and alpha^2) from a DevChannel in HBM instead of `alpha`.
void launch_fold_line(Ctx& c, const uint32_t* d_src, uint32_t log_src, qm31 alpha, uint32_t* d_dst,
                      uint8_t* d_leaf_out, const DevChannel* d_alpha = nullptr);
// `coords` is a host array of 4 device pointers.
void launch_fold_circle_into_line(Ctx& c, const uint32_t* const* coords, uint32_t log_src, qm31 alpha,
                                  uint32_t* d_dst, bool accumulate, uint8_t* d_leaf_out,
                                  const DevChannel* d_alpha = nullptr, uint32_t batch = 1, uint64_t bstride = 0);
// digest <- H(digest || root); alpha <- drawSecureFelt; root copied to d_root_out (optional).
void launch_chan_mix_root_draw(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, uint8_t* d_root_out);
// One tree reduction per launch from level k0 (tree_layer_kernel, kernels_fri.hip), for
// k0 <= TREE_LAYER_MAX_LOG.  Larger trees (and FRI layers) reduce their wide levels with
// merkle_parents launches first.  The launches share the context's completion counters,
// which every launch leaves at zero.
constexpr uint32_t TREE_LAYER_MAX_LOG = 16;
constexpr uint32_t TREE_CNT_GROUPS = 16;  // completion sub-counters, then the top counter
constexpr uint32_t TREE_CNT_STRIDE = 32;  // words between counters (a cache line each)
constexpr uint32_t TREE_XR_OFFSET = TREE_CNT_STRIDE * (TREE_CNT_GROUPS + 1);  // root exchange after them
constexpr uint32_t TREE_MAX_BLOCKS = 256;
constexpr uint32_t TREE_CNT_WORDS = TREE_XR_OFFSET + 8 * TREE_MAX_BLOCKS;  // counters + exchange of one tree launch
enum FriSrc { FRI_SRC_HASHES = 0, FRI_SRC_LINE = 1, FRI_SRC_CIRCLE = 2 };
struct FriLayerArgs {
    const uint4* src_line;      // FRI_SRC_LINE: the previous line (2^(k0+1) QM31 values)
    const uint32_t* coord[4];   // FRI_SRC_CIRCLE: the 4 coordinate columns (2^(k0+1) words each)
    const uint32_t* inv;        // inverse x (line) or y (circle) of the fold, 2^k0 words
    uint4* dst_line;            // the folded line (2^k0 values)
    uint8_t* tree;              // DeviceTree::layers
    uint32_t k0, s;             // level of the first hashes; log2 hashes per workgroup
    DevChannel* ch;             // channel step on the root (required by the folds)
    uint32_t* root_out;         // root copy (optional)
    uint32_t* cnt;              // completion counters (set by the launcher)
    uint32_t* xr;               // subtree root exchange, 8 words per workgroup (set by the launcher)
    B2State leaf_seed, node_seed;  // set by the launcher
    uint64_t bstride;           // batch: proof blockIdx.y's buffers bstride bytes after proof 0's,
                                // its counters and exchange TREE_CNT_WORDS words after
};
// Levels k0-1 .. 0 of `t` from its hashes at level k0; with d_ch, the channel step on the root.
void launch_tree_top(Ctx& c, DeviceTree& t, uint32_t k0, DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch = 1,
                     uint64_t bstride = 0, uint32_t* d_cnt = nullptr);
// One FRI inner layer: fold the line of log_src (or the 4 circle coordinates of log_src) with
// the alpha in d_ch into d_dst, commit it into `t` (max_log = log_src - 1) and run the channel
// step on its root (root copied to d_root_out).
void launch_fri_commit_line(Ctx& c, const uint32_t* d_src, uint32_t log_src, uint32_t* d_dst, DeviceTree& t,
                            DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch = 1, uint64_t bstride = 0,
                            uint32_t* d_cnt = nullptr);
void launch_fri_commit_circle(Ctx& c, const uint32_t* const* coords, uint32_t log_src, uint32_t* d_dst,
                              DeviceTree& t, DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch = 1,
                              uint64_t bstride = 0, uint32_t* d_cnt = nullptr);
// FRI inner layers below the committed line top_log (<= FRI_TAIL_MAX_LOG + 1, its alpha in
// ch) in one workgroup.  Out: lines top_log-1 .. last_log; for each log l in
// (last_log, top_log) the tree (every layer), its root copy (roots[l]) and the channel step.
constexpr uint32_t FRI_TAIL_MAX_LOG = 10;
struct FriTailArgs {
    uint32_t top_log, last_log;
    uint8_t* trees[FRI_TAIL_MAX_LOG + 2];  // DeviceTree::layers per log (last_log+1 .. top_log-1)
    uint4* lines[FRI_TAIL_MAX_LOG + 2];    // line evaluations per log (last_log .. top_log)
    uint32_t* roots[FRI_TAIL_MAX_LOG + 2]; // root copies per log
    DevChannel* ch;
    const uint32_t* itw_half;  // inverse tower + 2^(M-1) (set by the launcher)
    B2State leaf_seed, node_seed;  // set by the launcher
    uint64_t bstride;              // batch: proof blockIdx.x's buffers bstride bytes after proof 0's
};
void launch_fri_tail(Ctx& c, FriTailArgs& a, uint32_t batch = 1);

// LogUp / GKR (kernels_logup.hip).  Layer kinds of gkr_circuit.zig:79-92.
enum GkrKind { GKR_GRAND_PRODUCT = 0, GKR_LOGUP_GENERIC = 1, GKR_LOGUP_MULTIPLICITIES = 2, GKR_LOGUP_SINGLES = 3 };
void launch_fill_u32(Ctx& c, uint32_t* d, uint64_t n, uint32_t v);
void launch_zero_u64(Ctx& c, unsigned long long* d, uint32_t n);
void launch_xor_trace(Ctx& c, uint32_t log_size, uint64_t seed, uint32_t* a, uint32_t* b, uint32_t* cc,
                      uint32_t* mult /* 2^16 words */);
// den = z - (a + b alpha + c alpha^2) (QM31 AoS); table rows (i >> 8, i & 255, xor) when a == nullptr.
void launch_logup_denominators(Ctx& c, const uint32_t* a, const uint32_t* b, const uint32_t* cc, uint64_t n, qm31 z,
                               qm31 alpha, uint32_t* den);
// Cumulative column S_i = sum_{j<=i} num_j / den_j (num_kind 0 QM31, 1 M31, 2 ones); returns the claimed sum
// (synchronises); throws on a zero denominator.
qm31 logup_cumulative(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n, uint32_t* sums);
// The same without a readback: the claimed sum goes to d_total[0..4) (16-byte aligned) and the
// zero-denominator flag to *d_flag.
void logup_cumulative_dev(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n, uint32_t* sums,
                          uint32_t* d_total, uint32_t* d_flag);
// The same as logup_cumulative with no allocation: the block sums, the total and the flag live
// in a fixed per-device scratch array (n <= LOGUP_NOALLOC_MAX_ROWS); calls that overlap on one
// device race on it.  logup_scratch_init (once per device, that device current; called by
// ctx_init) looks up the array's device address.
constexpr int MAX_DEVICES = 64;
constexpr uint64_t LOGUP_NOALLOC_MAX_ROWS = 1ull << 27;
void logup_scratch_init(int device);
qm31 logup_cumulative_noalloc(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n,
                              uint32_t* sums);
// Denominators with z = d_z_alpha[0..4) and alpha = d_z_alpha[4..8) read on the device.
void launch_logup_denominators_dev(Ctx& c, const uint32_t* a, const uint32_t* b, const uint32_t* cc, uint64_t n,
                                   const uint32_t* d_z_alpha, uint32_t* den);
void launch_gkr_eq_evals(Ctx& c, const uint32_t* d_y, uint32_t n, qm31 scale, uint32_t* d_out);
void launch_gkr_next_layer(Ctx& c, int kind, const uint32_t* d_num, const uint32_t* d_den, uint32_t n_vars,
                           uint32_t* d_out_num, uint32_t* d_out_den);
// Adds the 64-bit coordinate sums of (f(0), f(2)) into d_acc8[0..8).
void launch_gkr_sums(Ctx& c, int kind, const uint32_t* d_eq, const uint32_t* d_num, const uint32_t* d_den,
                     uint32_t n_vars, qm31 lambda, unsigned long long* d_acc8);
void launch_mle_fix_first(Ctx& c, bool base, const uint32_t* d_in, uint32_t n_vars, qm31 r, uint32_t* d_out);
void launch_logup_table_cols(Ctx& c, const uint32_t* mult16, uint64_t n, uint32_t* mpad, uint32_t* negm);
struct GkrInstance {
    int kind;
    const uint32_t* d_num;  // QM31 AoS (generic), M31 (multiplicities) or null
    const uint32_t* d_den;  // QM31 AoS (values for grand product)
    uint32_t n_vars;
};
// Quotients (row_evaluation.zig).
// One (column, coefficient) term of the quotient row sums.
struct QEntryDev {
    const uint32_t* col;
    uint32_t shift;  // lift: row -> ((row >> shift) << 1) + (row & 1); 1 = same size
    uint32_t pad;
    uint32_t coeff[4];
};
static_assert(sizeof(QEntryDev) == 32, "quotient entry layout");
struct QuotientBatchDev {
    cm31 det, pix, piy;  // sample point components
    qm31 sum_a, sum_b;   // linear term
    uint32_t first, count;  // range into the (column, coeff) entry arrays
};
// Host arrays in: column device pointers and log sizes, per-entry column index
// and 4 coefficient words, batches (ranges into the entries).
void launch_quotients(Ctx& c, const uint32_t* const* h_cols, const uint32_t* h_col_logs, const uint32_t* h_entry_cols,
                      const uint32_t* h_entry_coeffs, uint32_t n_entries, const QuotientBatchDev* h_batches,
                      uint32_t n_batches, uint32_t log_size, uint32_t* const* out_coords /* host array of 4 */,
                      uint8_t* d_leaf_out);
// The same with the entries and batches already in device memory (in_bytes: algorithmic
// bytes read, for the profile only).
//
// Batches of proofs (the launchers' `batch`, `bstride` arguments): `batch` proofs of one shape in
// one launch (grid dimension y, or z where y is taken), proof k's buffers -- every per-proof
// pointer argument and every column pointer in a table -- at k * bstride bytes past proof 0's.
// Shared read-only data (twiddles, domain tables, seeds) is not offset.
// ident: every entry's shift is 1 (each column has the domain's size).
void launch_quotients_dev(Ctx& c, const QEntryDev* d_ent, uint32_t n_entries, const QuotientBatchDev* d_bat,
                          uint32_t n_batches, double in_bytes, uint32_t log_size, uint32_t* const* out_coords,
                          uint8_t* d_leaf_out, uint32_t batch = 1, uint64_t bstride = 0, bool ident = false);
// Point evaluation: sum_i coeff_i * w_i with weights of `point` (2^log).
void launch_eval_weights(Ctx& c, const qm31 factors[32], uint32_t log_size, uint32_t* d_w);
// The same with the log_size factors (QM31) in device memory.
// d_tab: the factor tables (eval_table_bytes, per proof of a batch), or NULL for a temporary.
void launch_eval_weights_dev(Ctx& c, const uint32_t* d_factors, uint32_t log_size, uint32_t* d_w, uint32_t batch = 1,
                             uint64_t bstride = 0, uint32_t* d_tab = nullptr);
void launch_eval_columns(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_size,
                         const uint32_t* d_w, uint32_t* d_out /* n_cols * 4 words */);
// The same two steps with caller-provided scratch: the factor tables (eval_table_bytes) and
// the per-chunk partial sums (eval_partial_bytes); nothing is allocated.
uint64_t eval_table_bytes(uint32_t log_size);
uint64_t eval_partial_bytes(uint32_t n_cols, uint32_t log_size);
void launch_eval_weights_into(Ctx& c, const qm31 factors[32], uint32_t log_size, uint32_t* d_w, uint32_t* d_tab);
void launch_eval_columns_into(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_size,
                              const uint32_t* d_w, uint32_t* d_out, unsigned long long* d_part, uint32_t batch = 1,
                              uint64_t bstride = 0);
// Wide Fibonacci.
// d_seeds (device, one per proof of a batch) replaces `seed` when non-NULL.
void launch_widefib_trace(Ctx& c, uint32_t log_n, uint32_t seq_len, uint32_t mode, uint64_t seed, uint32_t* d_out,
                          uint64_t col_stride, uint32_t batch = 1, const uint64_t* d_seeds = nullptr,
                          uint64_t bstride = 0);
void launch_widefib_composition(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_rows_eval,
                                uint32_t log_n_rows, const uint32_t* d_powers, uint32_t den_inv0,
                                uint32_t den_inv1, uint32_t* const* out_coords /* host array of 4 */,
                                uint32_t batch = 1, uint64_t bstride = 0);
// Poseidon AIR trace: 1264 columns of 2^log_n rows at d_out + column * col_stride.
void launch_poseidon_trace(Ctx& c, uint32_t log_n, uint32_t* d_out, uint64_t col_stride);
// Periodic XOR example AIR: main column and the IsFirst / IsStepWithOffset indicators.
void launch_xor_example_columns(Ctx& c, uint32_t log_size, uint32_t log_step, uint64_t offset, uint32_t* d_main,
                                uint32_t* d_is_first, uint32_t* d_is_step);
// Composition evaluation that is one QM31 constant on 2^log_eval rows.
void launch_constant_composition(Ctx& c, uint32_t log_eval, const uint32_t v[4], uint32_t* const* out_coords);
// PoW: lowest nonce in [base, base+count) or UINT64_MAX.
uint64_t launch_pow_grind(Ctx& c, const uint8_t prefix[32], uint32_t bits, uint64_t base, uint64_t count);
// The same for `batch` independent prefixes (d_prefix: 8 words each) over one nonce range:
// d_best[k] (initialised by the caller to UINT64_MAX) receives prefix k's lowest valid nonce.
void launch_pow_grind_batch(Ctx& c, const uint32_t* d_prefix, uint32_t batch, uint32_t bits, uint64_t base,
                            uint64_t count, unsigned long long* d_best);
// Gather words: out[i] = word at absolute device address d_addr[i].
void launch_gather(Ctx& c, const uint64_t* d_addr, uint64_t n, uint32_t* d_out);
void launch_tower(Ctx& c, uint32_t log, uint32_t* d_tw, uint32_t* d_itw, uint32_t* d_tw2, uint32_t* d_itw2);
void launch_domain_points(Ctx& c, uint32_t log, uint32_t* d_xy);
void launch_circle_inv_y(Ctx& c, uint32_t log, uint32_t* d_out);
void launch_batch_inverse(Ctx& c, const uint32_t* d_in, uint64_t n, uint32_t* d_out);
void launch_bit_reverse(Ctx& c, uint32_t* d_vals, uint32_t log_size, uint32_t words_per_elem);
void launch_accumulate(Ctx& c, uint32_t* d_acc, const uint32_t* d_src, uint64_t n_words);
void launch_qm31_to_coords(Ctx& c, const uint32_t* d_qm31, uint64_t n, uint32_t* const* coords /* host array */);

// Merkle commit of arbitrary columns (device pointers with log sizes); returns a resident tree.
DeviceTree* merkle_commit(Ctx& c, const std::vector<const uint32_t*>& cols, const std::vector<uint32_t>& logs);
DeviceTree* merkle_commit_qm31(Ctx& c, const uint32_t* d_qm31, uint32_t log_size);
void tree_read_root(Ctx& c, DeviceTree& t);
void tree_free(Ctx& c, DeviceTree* t);

// Upload a small host array into an arena buffer.
template <class T>
DBuf upload(Ctx& c, const T* host, size_t n) {
    DBuf b(&c, n * sizeof(T));
    if (n) SR_HIP(hipMemcpyAsync(b.p, c.stage(host, n * sizeof(T)), n * sizeof(T), hipMemcpyHostToDevice, c.stream));
    return b;
}

}  // namespace sr
