// Host-side proving driver of the rocm backend (internal; the C-ABI in
// include/stwo_rocm.h exposes it as stwo_rocm_prove_wide_fibonacci).
#pragma once
#include <functional>
#include <string>

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

struct ProveParams {
    uint32_t log_n_rows;
    uint32_t sequence_len;
    uint32_t pow_bits;
    uint32_t log_blowup_factor;
    uint32_t log_last_layer_degree_bound;
    uint32_t n_queries;
    uint32_t trace_mode;
    uint64_t trace_seed;
};

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

}  // namespace sr
