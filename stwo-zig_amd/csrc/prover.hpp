// Host-side proving driver of the rocm backend (internal; the C-ABI in
// include/stwo_rocm.h exposes it as stwo_rocm_prove_wide_fibonacci).
#pragma once
#include <algorithm>
#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <functional>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace sr {

// Blake2s Fiat-Shamir channel (src/core/channel/blake2s.zig:32-287).
struct Channel {
    uint8_t digest[32] = {0};
    uint32_t n_draws = 0;
    void mix_bytes(const void* data, size_t n);
    void mix_felts(const qm31* f, size_t n);
    void mix_u32s(const uint32_t* d, size_t n);
    void mix_u64(uint64_t v);
    void mix_root(const uint8_t root[32]);
    void draw_u32s(uint32_t out[8]);
    void draw_base_felts(uint32_t out[8]);
    qm31 draw_secure_felt();
    void pow_prefix(uint32_t bits, uint8_t out[32]) const;
};

// Owns the device Merkle trees of one proof (freed with it).
struct TreeOwner {
    Ctx& c;
    std::vector<DeviceTree*> trees;
    explicit TreeOwner(Ctx& cc) : c(cc) {}
    ~TreeOwner() {
        for (auto t : trees) tree_free(c, t);
    }
    DeviceTree* add(DeviceTree* t) {
        trees.push_back(t);
        return t;
    }
    DeviceTree* make(uint32_t max_log) {
        DeviceTree* t = new DeviceTree();
        t->max_log = max_log;
        trees.push_back(t);
        t->layers = (uint8_t*)c.alloc(((2ull << max_log) - 1) * 32);
        return t;
    }
};
// Committed tree: LDE columns, retained coefficients and the resident Merkle tree.
struct Committed {
    std::vector<uint32_t*> lde;
    std::vector<uint32_t> lde_logs;
    std::vector<const uint32_t*> coeffs;
    std::vector<uint32_t> coeff_logs;
    DBuf lde_buf;
    DeviceTree* tree = nullptr;
};

// Host-side image of the small device tables of one pipeline run (column pointer lists,
// leaf plans, quotient entries, an initial channel): built before the first launch and
// uploaded with one copy, each table at a 16-byte aligned offset.
struct HostPack {
    std::vector<uint8_t> buf;
    template <typename T>
    size_t add(const T* p, size_t n) {
        const size_t at = (buf.size() + 15) & ~(size_t)15;
        buf.resize(at + n * sizeof(T));
        if (n) memcpy(buf.data() + at, p, n * sizeof(T));
        return at;
    }
    size_t reserve(size_t bytes) {
        const size_t at = (buf.size() + 15) & ~(size_t)15;
        buf.resize(at + bytes, 0);
        return at;
    }
};
// commit_coefficients in two halves around the one upload: plan_commit allocates the LDE
// block and the tree and records the pointer tables and the leaf plan in the pack;
// launch_commit runs the LDE and the Merkle commitment from the uploaded copies and writes
// the root to d_root_out (optional).  Columns of one log size.
struct CommitPlan {
    uint32_t n = 0, m = 0, coeff_log = 0;
    size_t at_src = 0, at_dst = 0, at_tab = 0;
    std::vector<LeafSeg> segs;
};
void plan_commit(Ctx& c, TreeOwner& owner, Committed& out, const std::vector<const uint32_t*>& coeffs,
                 uint32_t coeff_log, uint32_t blowup, HostPack& pack, CommitPlan& cp);
void launch_commit(Ctx& c, Committed& out, const CommitPlan& cp, uint8_t* d_pack, uint8_t* d_root_out);
// The same from evaluations (the pack holds their pointer list at at_evals): interpolated into
// the plan's coefficient columns and extended, fused into one launch sequence at blowup 2.
void launch_commit_from_evals(Ctx& c, Committed& out, const CommitPlan& cp, uint8_t* d_pack, size_t at_evals,
                              uint8_t* d_root_out);

// LDE of coefficient columns (pointer list) by `blowup` and Merkle commit
// (commit_polys.zig:10-67; preparation.zig:113-210).
void commit_coefficients(Ctx& c, TreeOwner& owner, Committed& out, const std::vector<const uint32_t*>& coeffs,
                         uint32_t coeff_log, uint32_t blowup);

// FRI inner layers below a committed line (prover/fri.zig:497-586; the whole-cascade hook
// commitFriLayers, metal/commit_backend.zig:636-800).  `d_top` (log top_log) is committed and
// its fold alpha is in d_ch.  Each layer of log top_log-1 .. last_log+1 is folded from the one
// above, committed into a tree made by `owner` (appended to `trees`, its root copied to
// d_roots + 32 k for the k-th of them) and its root mixed into d_ch with the next alpha
// drawn; the line of log last_log is folded into d_last.  dst_lines[k] receives the line of
// log top_log-1-k (k < top_log-1-last_log).
void fri_fold_commit_down(Ctx& c, TreeOwner& owner, const uint32_t* d_top, uint32_t top_log, uint32_t last_log,
                          DevChannel* d_ch, uint32_t* const* dst_lines, uint32_t* d_last, uint8_t* d_roots,
                          std::vector<DeviceTree*>& trees);

// Quotient batches (core/pcs/quotients/samples.zig:245-312): one batch per distinct sample point
// in first-seen order over columns in tree/column/sample order, a two-point column adding its
// periodicity sample first; random coefficients 1, alpha, alpha^2, ... in that order.
struct SamplePoint {
    qm31 x, y;
};
struct ColumnSample {
    SamplePoint p;
    qm31 v;
};
struct QuotientPlan {
    std::vector<uint32_t> entry_cols, entry_coeffs;  // per entry: column index, 4 coefficient words
    std::vector<QuotientBatchDev> batches;
};
QuotientPlan plan_quotients(Ctx& c, uint32_t L, const std::vector<uint32_t>& col_logs,
                            const std::vector<std::vector<ColumnSample>>& samples, qm31 alpha);

// The end of a proof once the FRI layers are committed (fri.zig:594-621, scheme.zig:733-800,
// interop/proof_wire.zig:19-76): the last layer polynomial from the last line (values as read
// back, bit-reversed), PoW, the queries, one gather of every queried value and witness, and
// the proof wire.  trees: the committed trees in commitment order (composition last);
// n_samples: mask points per column (tree/column order); sv: the sampled values (4 words each,
// tree/column/sample order); the first FRI layer's tree and coordinate columns, the inner
// layers' trees and lines (log L-1 first).  `mark` is called after the FRI, PoW, decommitment
// and encoding steps.
struct ProofTail {
    uint32_t L = 0, last_log = 0;
    std::vector<const Committed*> trees;
    std::vector<uint32_t> n_samples;
    std::vector<uint32_t> sv;
    DeviceTree* ftree0 = nullptr;
    const uint32_t* qc[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<DeviceTree*> ftrees;
    std::vector<const uint32_t*> lines;
    std::vector<qm31> last_values;
};
struct ProveParams;
void finish_proof(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, std::string& proof,
                  const std::function<void()>& mark);

// Gather plan: absolute device word addresses, downloaded with one kernel.
struct Gather {
    std::vector<uint64_t> addr;
    size_t add_words(const uint32_t* p, uint64_t n) {
        size_t at = addr.size();
        for (uint64_t i = 0; i < n; i++) addr.push_back((uint64_t)(uintptr_t)(p + i));
        return at;
    }
    size_t add_hash(const DeviceTree& t, uint32_t layer, uint64_t idx) {
        return add_words((const uint32_t*)(t.layer(layer) + idx * 32), 8);
    }
};
// The phases of finish_proof: the last layer and its mix; with fp.nonce set, the nonce mix, the
// queries and the gather plan (appended to g); the gather; the wire encoding from the words.
struct FinishPlan {
    struct LayerPlan {
        std::vector<size_t> wit_at, hw_at;
    };
    std::vector<qm31> lastp;
    uint64_t nonce = 0;
    std::vector<uint64_t> qs;
    std::vector<size_t> qv_at;
    std::vector<std::vector<size_t>> tw_at;
    std::vector<LayerPlan> lplans;
};
void finish_last_layer(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, FinishPlan& fp);
void finish_queries(Ctx& c, Channel& ch, const ProveParams& prm, ProofTail& t, FinishPlan& fp, Gather& g);
void gather_words(Ctx& c, const Gather& g, std::vector<uint32_t>& gw);
std::string finish_encode(const ProveParams& prm, const ProofTail& t, const FinishPlan& fp, const uint32_t* gw);

// Components served by the one-trace-tree proof core: the wide-Fibonacci AIR and the Poseidon
// AIR (examples/poseidon.zig:397-504: 1264 columns, a constant composition evaluation).
// The periodic XOR example AIR (examples/xor.zig:419-605) adds a preprocessed tree of two
// indicator columns to one main column.
enum : uint32_t { AIR_WIDE_FIBONACCI = 0, AIR_POSEIDON = 1, AIR_XOR = 2 };
constexpr uint32_t POSEIDON_COLUMNS = 1264, POSEIDON_COLUMNS_PER_REP = 158;

struct ProveParams {
    uint32_t log_n_rows;
    uint32_t sequence_len;
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
    uint32_t trace_mode;
    uint64_t trace_seed;
    uint32_t air = AIR_WIDE_FIBONACCI;  // the component proven (sequence_len is its column count)
    uint32_t xor_log_step = 0;          // XOR statement (log_size is log_n_rows)
    uint64_t xor_offset = 0;
};

void trace_vanishing_inverses(Ctx& c, uint32_t n, uint32_t den_inv[2]);

// Runs f(0) .. f(n-1) on the process's host worker pool (up to `threads` - 1 workers plus the
// calling thread) and returns true, or returns false without running anything when the pool is
// taken by another caller (a nested call or another thread), who then runs its loop itself.
// The first exception is stored in err after every index has finished or been skipped.
bool host_pool_run(uint32_t n, uint32_t threads, const std::function<void(uint32_t)>& f, std::exception_ptr& err);

// f(0) .. f(n-1) on up to 16 host threads (the proofs' host phases are independent): the
// workers live for the process, so a call costs a wake-up rather than thread start-ups; the
// first exception is rethrown after every index has finished or been skipped.
template <class F>
void parallel_for(uint32_t n, F&& f) {
    const uint32_t hw = std::max(1u, std::thread::hardware_concurrency());
    const uint32_t T = std::min(std::min(hw, 16u), n);
    if (T > 1) {
        std::exception_ptr err;
        const std::function<void(uint32_t)> fn = [&f](uint32_t i) { f(i); };
        if (host_pool_run(n, T, fn, err)) {
            if (err) std::rethrow_exception(err);
            return;
        }
    }
    for (uint32_t i = 0; i < n; i++) f(i);
}


// Many wide-Fibonacci proofs of one shape and PCS configuration in single launches
// (batch_prover.hip): proof k uses prm with trace_seed = seeds[k] (seeds may be NULL: every
// proof uses prm.trace_seed).  Requires n + log_blowup <= TREE_LAYER_MAX_LOG + 1 and
// log_last_layer_degree_bound + log_blowup <= FRI_TAIL_MAX_LOG.
struct BatchStats {
    double device_ms = 0;  // first launch to the readback of the FRI state
    double tail_ms = 0;    // last layers, proof of work, gather and encoding of every proof
    double total_ms = 0;
    uint64_t slice_bytes = 0;  // device bytes per proof
    // host wall time of the phases: set-up (tables, uploads, launches), wait for the device,
    // last layers, proof of work, query plans and gather, wire encoding
    double phase_ms[6] = {0, 0, 0, 0, 0, 0};
};
void prove_wide_fibonacci_batch(Ctx& c, const ProveParams& prm, uint32_t n_proofs, const uint64_t* seeds,
                                std::vector<std::string>& proofs, BatchStats* stats);
// The same in two steps, for repeated batches of one shape: a plan holds the device allocation,
// tables and pinned readback of n_proofs proofs; each run proves one batch of seeds.  With
// use_graph the device part of the first run is recorded into a hipGraph and later runs
// replay it (a plan whose quotients need a temporary buffer, or a profiled context, runs
// eagerly).  Plans are tied to their context and freed before it.
struct BatchPlan;
BatchPlan* batch_plan_create(Ctx& c, const ProveParams& prm, uint32_t n_proofs);
void batch_plan_run(Ctx& c, BatchPlan& plan, const uint64_t* seeds, bool use_graph, std::vector<std::string>& proofs,
                    BatchStats* stats);
void batch_plan_free(BatchPlan* plan);
uint64_t batch_plan_slice_bytes(const BatchPlan& plan);

constexpr int STAGES = 8;
struct ProveStats {
    // main_commit, composition, sampled_values, quotients, fri_commit, pow, decommit, encode
    double stage_gpu_ms[STAGES];
    double stage_host_ms[STAGES];
    double total_gpu_ms;
    double total_host_ms;
    uint64_t peak_hbm_bytes;
};

// d_trace: W columns of 2^n words, column-major with stride 2^n (NULL: generate
// with ProveParams.trace_mode / trace_seed inside the proof).
void prove_wide_fibonacci(Ctx& c, const ProveParams& prm, const uint32_t* d_trace, std::string& proof,
                          ProveStats* stats);

// The XOR-lookup AIR as a whole proof (logup_proof.hip): trace from (log_n, seed), PCS
// configuration from prm (pow_bits, log_blowup_factor, log_last_layer_degree_bound, n_queries).
void prove_logup_xor(Ctx& c, uint32_t log_n, uint64_t seed, const ProveParams& prm, std::string& proof);

// GKR batch proof of the given input layers on `ch` (device-resident layers);
// returns the proof word stream described in gkr.hip.
std::vector<uint32_t> gkr_prove_batch(Ctx& c, Channel& ch, const std::vector<GkrInstance>& inst);
// The same on a device channel, without a host round trip: gkr_launch_on_channel enqueues the
// whole proof (the state starts from and ends in d_ch); gkr_collect reads the proof back
// (one synchronisation) and returns the word stream.
struct GkrRun;
GkrRun* gkr_launch_on_channel(Ctx& c, DevChannel* d_ch, const std::vector<GkrInstance>& inst);
std::vector<uint32_t> gkr_collect(Ctx& c, GkrRun* run);
void gkr_release(GkrRun* run);
// One sum-check round polynomial of a layer (eq table on the device, y = eq point).
std::vector<qm31> gkr_sum_as_poly(Ctx& c, int kind, const uint32_t* d_eq, const uint32_t* d_num,
                                  const uint32_t* d_den, uint32_t n_vars, qm31 lambda, qm31 claim,
                                  const std::vector<qm31>& y, qm31 correction);

// LogUp workload (XOR-lookup AIR) from trace to quotient commitment; returns the word
// stream described in logup_pipeline.hip.  total_ms (optional): host wall time.
std::vector<uint32_t> logup_xor_pipeline(Ctx& c, uint32_t log_n, uint64_t seed, double* total_ms);

// Batched primitive pipeline (batched_pipeline.hip): LDE + lifted Merkle of many
// columns and FRI fold_line rounds, planned once and replayed as one hipGraph.
struct BatchedSpec {
    uint32_t n_columns, log_min, log_max, log_blowup, fri_log, n_folds;
    uint64_t seed;
};
uint32_t batched_column_log(const BatchedSpec& s, uint32_t i);
constexpr size_t CH_BYTES = 128;  // DevChannel slot ahead of the root copies
static_assert(sizeof(DevChannel) <= CH_BYTES, "DevChannel slot");
class BatchedPipeline {
   public:
    struct Tuning {
        uint32_t fft_tile_log, merkle_levels;
        double ms;
    };
    BatchedPipeline(Ctx& c, const BatchedSpec& s);
    ~BatchedPipeline();
    void record();                  // the launch sequence (eager, or under capture), with this pipeline's tuning
    void capture();                 // capture record() into the graph
    double replay(uint32_t n);      // n graph launches (recaptured first if the tower moved); ms per launch
    std::vector<uint32_t> result();  // roots, final channel digest and draw counter
    void autotune(uint32_t reps);    // pick fft_tile_log x merkle_levels, recapture
    void set_tuning(uint32_t fft_tile_log, uint32_t merkle_levels);  // recaptures
    std::vector<Tuning> tuning;
    uint32_t tile_log = 13, levels = 10;  // this pipeline's launch shapes (the context's are left alone)
    uint64_t input_words() const;

   private:
    // The twiddle tower the graph was captured against: a later ensure_tower on the context
    // (a larger proof, twiddles_prepare) moves it and returns the old arrays to the arena
    // cache, so replay recaptures when these differ from the context's.
    const uint32_t* cap_tower[4] = {nullptr, nullptr, nullptr, nullptr};
    uint32_t cap_tower_log = 0;
    bool tower_moved() const;
    struct SizeClass {
        uint32_t log = 0, n = 0;
        DBuf d_eval, d_coeff, d_lde;
    };
    Ctx& c;
    BatchedSpec sp;
    TreeOwner owner;
    std::vector<uint32_t> logs;
    uint32_t max_lde = 0;
    std::vector<LeafSeg> leaf_segs;  // leaf-hash segment plan of the lifted tree
    DBuf evals, coeffs, lde, leaf_tab, chan;
    std::vector<SizeClass> classes;
    DeviceTree* tree = nullptr;
    std::vector<DBuf> lines;
    std::vector<DeviceTree*> ftrees;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Transcript steps on the device channel between commitments (transcript.hip).
// mixRoot(root) [+ mixU32s(words[0..n_words))], alpha = drawSecureFelt, powers[k] = alpha^k.
// (batch, bstride: one workgroup per proof, its buffers bstride bytes apart; see launch_quotients_dev)
void launch_tr_commit_draw(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, const uint32_t* words, uint32_t n_words,
                           uint32_t* d_powers, uint32_t n_pow, uint32_t batch = 1, uint64_t bstride = 0);
// digest <- H(digest || words[0..n_words)) if n_words > 0 or mix_empty (words in device
// memory), then n_draws secure felts into d_out (4 words each).
void launch_tr_mix_draws(Ctx& c, DevChannel* d_ch, const uint32_t* d_words, uint32_t n_words, uint32_t n_draws,
                         uint32_t* d_out, bool mix_empty = false);
// [mixRoot(root) if d_root], t = drawSecureFelt, OODS point (x, y) and the point-evaluation factors.
// n_blocks draws of 8 raw words each (drawU32s) into d_out (8 n_blocks words).
void launch_tr_draw_words(Ctx& c, DevChannel* d_ch, uint32_t n_blocks, uint32_t* d_out);
void launch_tr_oods(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, uint32_t n_factors, uint32_t* d_oods,
                    uint32_t* d_factors, uint32_t batch = 1, uint64_t bstride = 0);
// mixFelts(sampled values), q = drawSecureFelt, coefficients of the entries (column order)
// and the single batch at the OODS point (entries' col/shift set by the caller).
void launch_tr_quotient(Ctx& c, DevChannel* d_ch, const uint32_t* d_sv, uint32_t ns, const uint32_t* d_oods,
                        QEntryDev* d_ents, QuotientBatchDev* d_bat, uint32_t batch = 1, uint64_t bstride = 0);

// Continuous-batching proof service (serving.hip).
struct ServeJob {
    uint32_t kind;  // 0 wide-Fibonacci proof (seeded uniform trace), 1 XOR-lookup LogUp workload
    uint32_t log_n, seq_len, reserved;
    uint64_t seed;
    double arrival_ms;  // relative to the start of serve(); non-decreasing
};
struct ServeOutcome {
    uint8_t digest[32];  // Blake2s-256 of the output bytes
    double arrival_ms, start_ms, end_ms;
    uint32_t slot, status;
    uint64_t output_bytes;
};
struct ServeStats {
    double makespan_ms, p50_ms, p99_ms, mean_ms, jobs_per_s;
    uint64_t peak_admitted_bytes;
    uint64_t peak_device_bytes;  // device bytes held by all slots (live + cached), peak of the call
    uint32_t max_in_flight, n_slots;
};
uint64_t serve_job_bytes(const ServeJob& j);  // admission estimate of a job's HBM footprint
// Slot contexts kept between serve() calls (streams, warm arenas, towers, pinned staging):
// a service keeps its resources resident instead of rebuilding them per batch of requests.
struct ServePool;
ServePool* serve_pool_new();
void serve_pool_free(ServePool* p);
void serve(ServePool& pool, int device, const std::vector<ServeJob>& jobs, uint32_t n_slots, uint64_t hbm_budget,
           std::vector<ServeOutcome>& out, ServeStats& st);

}  // namespace sr
