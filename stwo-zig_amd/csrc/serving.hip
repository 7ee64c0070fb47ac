// Continuous-batching proof service on one GPU.
//
// The reference admits one job at a time against a resource profile
// (src/prover/native/resource_admission.zig) and proves it; here a request loop
// keeps n_slots proof slots busy: every slot owns a device context (its own HIP
// stream, caching arena, twiddle tower and resident Merkle layers; the slots persist in
// the library context between calls, so arenas stay warm), requests
// become eligible at their arrival time, and the head of the FIFO is admitted as
// soon as a slot is idle and its estimated HBM footprint fits the shared budget.
// Jobs of different slots overlap on the device (independent streams).  Every
// result is reduced to a Blake2s-256 digest of its output bytes, so a job's result
// can be compared alone and in the mix.
//
// Job kinds: 0 = wide-Fibonacci proof (functional protocol, seeded uniform trace,
// examples/wide_fibonacci.zig:181-343), 1 = XOR-lookup LogUp workload
// (logup_pipeline.hip), 2 = Poseidon AIR proof, 3 = periodic XOR AIR proof, 4 = XOR-lookup AIR
// proof (logup_proof.hip).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "prover.hpp"

namespace sr {

uint64_t serve_job_bytes(const ServeJob& j) {
    if (j.kind == 2) {  // Poseidon: 1264 columns of 2^(log_n_instances - 3) rows
        const uint64_t R = 1ull << (j.log_n - 3);
        return 16ull * R * POSEIDON_COLUMNS + 1024ull * R + (64ull << 20);
    }
    const uint64_t N = 1ull << j.log_n;
    if (j.kind == 0) return 16ull * N * j.seq_len + 1024ull * N + (64ull << 20);
    if (j.kind == 4) return 1536ull * N + (64ull << 20);  // 4 trees of 4..8 columns, pair fractions
    if (j.kind == 3) return 48ull * N + 1024ull * N + (64ull << 20);  // 3 columns
    return 1024ull * N + (64ull << 20);
}

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

struct Slot {
    Ctx c;
    ~Slot() { ctx_release(c); }
};

}  // namespace

struct ServePool {
    MemLedger ledger;  // device bytes held by all slots (declared first: outlives the slots)
    std::vector<std::unique_ptr<Slot>> slots;
};
ServePool* serve_pool_new() { return new ServePool(); }
void serve_pool_free(ServePool* p) { delete p; }

void serve(ServePool& pool, int device, const std::vector<ServeJob>& jobs, uint32_t n_slots, uint64_t hbm_budget,
           std::vector<ServeOutcome>& out, ServeStats& st) {
    SR_CHECK(n_slots >= 1 && n_slots <= 16, "serve: n_slots must be 1..16");
    SR_CHECK(!jobs.empty(), "serve: no jobs");
    uint32_t max_log = 0;
    for (size_t i = 0; i < jobs.size(); i++) {
        const ServeJob& j = jobs[i];
        SR_CHECK(j.kind <= 4, "serve: unknown job kind");
        SR_CHECK(j.kind != 4 || (j.log_n >= 16 && j.log_n <= 24), "serve: XOR-lookup AIR shape");
        SR_CHECK(j.kind != 3 || (j.log_n >= 2 && j.log_n <= 24 && j.seq_len <= j.log_n), "serve: XOR shape");
        SR_CHECK(j.kind != 2 || (j.log_n >= 4 && j.log_n <= 24), "serve: Poseidon shape");
        SR_CHECK(j.kind != 0 || (j.log_n >= 5 && j.log_n <= 24 && j.seq_len >= 3 && j.seq_len <= 256),
                 "serve: wide-Fibonacci shape");
        SR_CHECK(j.kind != 1 || (j.log_n >= 16 && j.log_n <= 24), "serve: LogUp shape");
        SR_CHECK(i == 0 || jobs[i - 1].arrival_ms <= j.arrival_ms, "serve: arrivals must be sorted");
        SR_CHECK(hbm_budget == 0 || serve_job_bytes(j) <= hbm_budget, "serve: a job exceeds the HBM budget");
        max_log = std::max(max_log, (j.kind == 2 ? j.log_n - 3 : j.log_n) + 2);
    }
    // Slots: resident contexts (kept in the pool between calls) with the twiddle tower and the
    // pinned staging ready, so no job pays set-up.
    std::vector<std::unique_ptr<Slot>>& slots = pool.slots;
    if (hbm_budget) {
        // Slots kept from earlier calls hold towers and domain tables of the sizes they served:
        // when those (after trimming the cached blocks) leave no room for this call's largest
        // job, the slots are rebuilt and hold only what this call needs.
        uint64_t max_job = 0, held = 0;
        for (const ServeJob& j : jobs) max_job = std::max(max_job, serve_job_bytes(j));
        for (auto& sl : slots) {
            sl->c.trim();
            held += sl->c.bytes_live + sl->c.bytes_cached;
        }
        if (held + max_job > hbm_budget) slots.clear();
    }
    while (slots.size() < n_slots) {
        std::unique_ptr<Slot> sl(new Slot());
        sl->c.ledger = &pool.ledger;
        ctx_init(sl->c, device, 0);
        uint8_t z = 0;
        sl->c.stage(&z, 1);
        (void)sl->c.pinned(1u << 20);
        slots.push_back(std::move(sl));
    }
    // Under a budget, slots left from an earlier call with more slots would hold memory that
    // this call does not count: release them.
    if (hbm_budget) slots.resize(n_slots);
    for (uint32_t s = 0; s < n_slots; s++) {
        slots[s]->c.ensure_tower(max_log);
        slots[s]->c.sync();
    }
    out.assign(jobs.size(), ServeOutcome{});
    std::mutex mu;
    std::condition_variable cv;
    std::deque<size_t> ready;  // arrived, not yet admitted (FIFO)
    uint64_t used = 0, peak = 0;
    uint32_t in_flight = 0, max_in_flight = 0;
    bool done_arriving = false;
    std::string first_error;
    // Device bytes each slot holds when idle (live: tower, counters, domain tables; cached arena
    // blocks), published at its idle points.  The budget covers what the slots really hold: a
    // job is admitted only if the idle holdings of every slot, the estimates of the jobs in
    // flight (each job can allocate at most its estimate beyond its slot's holdings) and its own
    // estimate fit; otherwise the admitting slot trims its cache, and running slots trim theirs
    // when their jobs end, until it fits.
    std::vector<uint64_t> idle_held(n_slots), idle_cached(n_slots);
    for (uint32_t s = 0; s < n_slots; s++) {
        idle_held[s] = slots[s]->c.bytes_live + slots[s]->c.bytes_cached;
        idle_cached[s] = slots[s]->c.bytes_cached;
    }
    bool trim_wanted = false;
    uint64_t version = 0;  // bumped on every change the admission depends on
    pool.ledger.reset_peak();
    const double t0 = now_ms();

    auto worker = [&](uint32_t sid) {
        Ctx& c = slots[sid]->c;
        if (hipSetDevice(device) != hipSuccess) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_error.empty()) first_error = "hipSetDevice failed in a serving slot";
            cv.notify_all();
            return;
        }
        for (;;) {
            size_t ji;
            {
                std::unique_lock<std::mutex> lk(mu);
                for (;;) {
                    cv.wait(lk, [&] { return !first_error.empty() || !ready.empty() || done_arriving; });
                    if (!first_error.empty() || ready.empty()) return;
                    const uint64_t need = serve_job_bytes(jobs[ready.front()]);
                    uint64_t idle = 0, cached = 0;
                    for (uint32_t t = 0; t < n_slots; t++) {
                        idle += idle_held[t];
                        cached += idle_cached[t];
                    }
                    if (hbm_budget == 0 || idle + used + need <= hbm_budget) break;
                    if (c.bytes_cached > 0) {
                        // this slot is idle: give its cached blocks back to the device
                        lk.unlock();
                        c.trim();
                        lk.lock();
                        idle_held[sid] = c.bytes_live;
                        idle_cached[sid] = 0;
                        version++;
                        continue;
                    }
                    if (in_flight == 0 && cached == 0) {
                        first_error = "serve: a job does not fit the HBM budget beside the slots' resident memory";
                        cv.notify_all();
                        return;
                    }
                    trim_wanted = true;
                    const uint64_t seen = version;
                    cv.wait(lk, [&] { return !first_error.empty() || version != seen; });
                }
                trim_wanted = false;
                ji = ready.front();
                ready.pop_front();
                used += serve_job_bytes(jobs[ji]);
                peak = std::max(peak, used);
                in_flight++;
                max_in_flight = std::max(max_in_flight, in_flight);
                version++;
            }
            ServeOutcome& o = out[ji];
            o.slot = sid;
            o.start_ms = now_ms() - t0;
            try {
                const ServeJob& j = jobs[ji];
                if (j.kind == 0) {
                    ProveParams prm{};
                    prm.log_n_rows = j.log_n;
                    prm.sequence_len = j.seq_len;
                    prm.pow_bits = 10;
                    prm.log_blowup_factor = 1;
                    prm.log_last_layer_degree_bound = 0;
                    prm.n_queries = 3;
                    prm.trace_mode = 1;
                    prm.trace_seed = j.seed;
                    std::string proof;
                    prove_wide_fibonacci(c, prm, nullptr, proof, nullptr);
                    b2_hash(proof.data(), proof.size(), o.digest);
                    o.output_bytes = proof.size();
                } else if (j.kind == 3) {
                    ProveParams prm{};
                    prm.log_n_rows = j.log_n;
                    prm.sequence_len = 1;
                    prm.pow_bits = 10;
                    prm.log_blowup_factor = 1;
                    prm.log_last_layer_degree_bound = 0;
                    prm.n_queries = 3;
                    prm.air = AIR_XOR;
                    prm.xor_log_step = j.seq_len;
                    prm.xor_offset = j.seed;
                    std::string proof;
                    prove_wide_fibonacci(c, prm, nullptr, proof, nullptr);
                    b2_hash(proof.data(), proof.size(), o.digest);
                    o.output_bytes = proof.size();
                } else if (j.kind == 2) {
                    ProveParams prm{};
                    prm.log_n_rows = j.log_n - 3;
                    prm.sequence_len = POSEIDON_COLUMNS;
                    prm.pow_bits = 10;
                    prm.log_blowup_factor = 1;
                    prm.log_last_layer_degree_bound = 0;
                    prm.n_queries = 3;
                    prm.air = AIR_POSEIDON;
                    std::string proof;
                    prove_wide_fibonacci(c, prm, nullptr, proof, nullptr);
                    b2_hash(proof.data(), proof.size(), o.digest);
                    o.output_bytes = proof.size();
                } else if (j.kind == 4) {
                    ProveParams prm{};
                    prm.pow_bits = 10;
                    prm.log_blowup_factor = 1;
                    prm.log_last_layer_degree_bound = 0;
                    prm.n_queries = 3;
                    std::string proof;
                    prove_logup_xor(c, j.log_n, j.seed, prm, proof);
                    b2_hash(proof.data(), proof.size(), o.digest);
                    o.output_bytes = proof.size();
                } else {
                    std::vector<uint32_t> w = logup_xor_pipeline(c, j.log_n, j.seed, nullptr);
                    b2_hash(w.data(), 4 * w.size(), o.digest);
                    o.output_bytes = 4 * w.size();
                }
                o.status = 0;
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_error.empty()) first_error = e.what();
                o.status = 1;
            }
            o.end_ms = now_ms() - t0;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (trim_wanted && c.bytes_cached > 0) {
                    lk.unlock();
                    c.trim();
                    lk.lock();
                }
                used -= serve_job_bytes(jobs[ji]);
                in_flight--;
                idle_held[sid] = c.bytes_live + c.bytes_cached;
                idle_cached[sid] = c.bytes_cached;
                version++;
            }
            cv.notify_all();
        }
    };
    std::vector<std::thread> th;
    for (uint32_t s = 0; s < n_slots; s++) th.emplace_back(worker, s);
    // Arrivals: each request becomes eligible at its arrival time.
    for (size_t i = 0; i < jobs.size(); i++) {
        const double wait = jobs[i].arrival_ms - (now_ms() - t0);
        if (wait > 0) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(wait * 1000.0)));
        {
            std::lock_guard<std::mutex> lk(mu);
            out[i].arrival_ms = jobs[i].arrival_ms;
            ready.push_back(i);
            if (i + 1 == jobs.size()) done_arriving = true;
            version++;
        }
        cv.notify_all();
    }
    for (auto& t : th) t.join();
    SR_CHECK(first_error.empty(), "serve: job failed: " + first_error);
    std::vector<double> lat(jobs.size());
    double last_end = 0, sum = 0;
    for (size_t i = 0; i < jobs.size(); i++) {
        lat[i] = out[i].end_ms - out[i].arrival_ms;
        sum += lat[i];
        last_end = std::max(last_end, out[i].end_ms);
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double q) {
        size_t k = (size_t)std::ceil(q * (double)lat.size()) - 1;
        return lat[std::min(k, lat.size() - 1)];
    };
    st.makespan_ms = last_end - jobs.front().arrival_ms;
    st.p50_ms = pct(0.50);
    st.p99_ms = pct(0.99);
    st.mean_ms = sum / (double)jobs.size();
    st.jobs_per_s = (double)jobs.size() / (st.makespan_ms * 1e-3);
    st.peak_admitted_bytes = peak;
    st.peak_device_bytes = pool.ledger.peak.load();
    st.max_in_flight = max_in_flight;
    st.n_slots = n_slots;
}

}  // namespace sr
