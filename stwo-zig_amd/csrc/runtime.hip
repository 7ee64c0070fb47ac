// Device context of the rocm backend: caching HBM arena (no hipMalloc on the
// steady-state proving path), pinned staging, resident twiddle tower and
// per-size domain tables.  Mirrors the role of the reference's Metal shared
// runtime (src/backends/metal/shared_runtime.zig:1-40) and twiddle tower
// session (src/prover/poly/twiddle_tower.zig:29-115, src/prover/session.zig:10-60).
#include "runtime.hpp"

namespace sr {

static size_t round_up(size_t b) {
    // 256-byte alignment; bucket large sizes to 1 MiB so buffers are reused across sizes.
    if (b <= (1u << 20)) return (b + 255) & ~(size_t)255;
    return (b + (1u << 20) - 1) & ~(size_t)((1u << 20) - 1);
}

void* Ctx::alloc(size_t bytes) {
    size_t sz = round_up(bytes ? bytes : 1);
    auto it = free_blocks.lower_bound(sz);
    // Reuse a cached block when it is not more than 25% oversized.
    if (it != free_blocks.end() && it->first <= sz + sz / 4) {
        void* p = it->second;
        size_t have = it->first;
        free_blocks.erase(it);
        bytes_cached -= have;
        live_blocks[p] = have;
        bytes_live += have;
        if (bytes_live > peak_live) peak_live = bytes_live;
        return p;
    }
    if (hbm_budget && bytes_live + sz > hbm_budget) {
        trim();
        if (bytes_live + sz > hbm_budget)
            throw Error(4, "HBM budget exceeded: live " + std::to_string(bytes_live) + " + request " +
                               std::to_string(sz) + " > budget " + std::to_string(hbm_budget));
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sz);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        trim();
        e = hipMalloc(&p, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            throw Error(2, "hipMalloc of " + std::to_string(sz) + " bytes failed");
        }
    }
    live_blocks[p] = sz;
    bytes_live += sz;
    if (bytes_live > peak_live) peak_live = bytes_live;
    if (ledger) ledger->add(sz);
    return p;
}

void Ctx::free(void* p) {
    if (!p) return;
    auto it = live_blocks.find(p);
    if (it == live_blocks.end()) return;
    size_t sz = it->second;
    live_blocks.erase(it);
    bytes_live -= sz;
    free_blocks.insert({sz, p});
    bytes_cached += sz;
}

void ctx_init(Ctx& c, int device, uint64_t hbm_budget) {
    int n = 0;
    SR_HIP(hipGetDeviceCount(&n));
    SR_CHECK(device >= 0 && device < n, "no such HIP device");
    c.device = device;
    SR_HIP(hipSetDevice(device));
    SR_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));  // default priority
    SR_HIP(hipEventCreateWithFlags(&c.ev_sync, hipEventDisableTiming));
    c.hbm_budget = hbm_budget;
    c.leaf_seed = b2_prefix_seed("leaf");
    c.node_seed = b2_prefix_seed("node");
    c.gen = make_gen_doubles();
    logup_scratch_init(device);
    const size_t cnt_bytes = 4ull * (TREE_XR_OFFSET + 8 * TREE_MAX_BLOCKS);
    c.d_tree_cnt = (uint32_t*)c.alloc(cnt_bytes);
    SR_HIP(hipMemsetAsync(c.d_tree_cnt, 0, cnt_bytes, c.stream));
    SR_HIP(hipStreamSynchronize(c.stream));
}

void ctx_release(Ctx& c) {
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    if (c.ev_sync) (void)hipEventDestroy(c.ev_sync);
    c.ev_sync = nullptr;
    for (auto& kv : c.live_blocks) (void)hipFree(kv.first);
    for (auto& kv : c.free_blocks) (void)hipFree(kv.second);
    c.live_blocks.clear();
    c.free_blocks.clear();
    if (c.ledger) c.ledger->sub(c.bytes_live + c.bytes_cached);
    c.bytes_live = c.bytes_cached = 0;
    if (c.h_pinned) (void)hipHostFree(c.h_pinned);
    if (c.h_ring) (void)hipHostFree(c.h_ring);
    c.h_pinned = nullptr;
    c.h_ring = nullptr;
    for (auto e : c.ev_pool) (void)hipEventDestroy(e);
    c.ev_pool.clear();
    if (c.stream) (void)hipStreamDestroy(c.stream);
    c.stream = nullptr;
    c.d_tw = c.d_itw = c.d_tw2 = c.d_itw2 = nullptr;
    c.d_tree_cnt = nullptr;
    c.tower_log = 0;
}

void Ctx::trim() {
    if (free_blocks.empty()) return;
    (void)hipStreamSynchronize(stream);
    for (auto& kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    if (ledger) ledger->sub(bytes_cached);
    bytes_cached = 0;
}

hipEvent_t Ctx::pool_event() {
    if (ev_used == ev_pool.size()) {
        hipEvent_t e;
        SR_HIP(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    return ev_pool[ev_used++];
}

// Waits for the stream by polling an event: a blocking wait may put the thread to sleep
// and add its wake-up latency to every transcript round trip.
void Ctx::sync() {
    SR_HIP(hipEventRecord(ev_sync, stream));
    hipError_t e;
    while ((e = hipEventQuery(ev_sync)) == hipErrorNotReady) {
    }
    SR_HIP(e);
    ring_off = 0;
}

uint8_t* Ctx::pinned(size_t bytes) {
    if (bytes > h_pinned_size) {
        SR_HIP(hipStreamSynchronize(stream));
        if (h_pinned) (void)hipHostFree(h_pinned);
        size_t sz = bytes < (1u << 20) ? (1u << 20) : bytes;
        SR_HIP(hipHostMalloc((void**)&h_pinned, sz, hipHostMallocDefault));
        h_pinned_size = sz;
    }
    return h_pinned;
}

void* Ctx::stage(const void* src, size_t bytes) {
    size_t need = (bytes + 255) & ~(size_t)255;
    if (need > ring_cap) {
        SR_HIP(hipStreamSynchronize(stream));
        if (h_ring) (void)hipHostFree(h_ring);
        ring_cap = need > (4u << 20) ? need : (4u << 20);
        SR_HIP(hipHostMalloc((void**)&h_ring, ring_cap, hipHostMallocDefault));
        ring_off = 0;
    }
    if (ring_off + need > ring_cap) {
        // Every earlier slot may still be read by a pending copy: drain the stream.
        SR_HIP(hipStreamSynchronize(stream));
        ring_off = 0;
    }
    uint8_t* dst = h_ring + ring_off;
    memcpy(dst, src, bytes);
    ring_off += need;
    return dst;
}

void Ctx::ensure_tower(uint32_t log) {
    if (log < 3) log = 3;
    if (tower_log >= log) return;
    SR_CHECK(log <= 30, "twiddle tower: circle log too large");
    uint64_t n = 1ull << (log - 1);
    uint32_t* tw = (uint32_t*)alloc(n * 4);
    uint32_t* itw = (uint32_t*)alloc(n * 4);
    uint32_t* tw2 = (uint32_t*)alloc(n * 4);
    uint32_t* itw2 = (uint32_t*)alloc(n * 4);
    launch_tower(*this, log, tw, itw, tw2, itw2);
    if (d_tw) {
        free(d_tw);
        free(d_itw);
        free(d_tw2);
        free(d_itw2);
    }
    d_tw = tw;
    d_itw = itw;
    d_tw2 = tw2;
    d_itw2 = itw2;
    tower_log = log;
    tower_builds++;
}

const uint32_t* Ctx::twiddle_layer(bool inverse, uint32_t m, uint32_t s) const {
    const uint32_t* t = inverse ? d_itw : d_tw;
    return t + ((1ull << (tower_log - 1)) - (1ull << (m - s)));
}

const uint32_t* Ctx::domain_points(uint32_t log) {
    auto it = domain_xy.find(log);
    if (it != domain_xy.end()) return it->second;
    uint32_t* p = (uint32_t*)alloc((2ull << log) * 4);
    launch_domain_points(*this, log, p);
    domain_xy[log] = p;
    return p;
}

const uint32_t* Ctx::inv_circle_y(uint32_t log) {
    auto it = circle_inv_y.find(log);
    if (it != circle_inv_y.end()) return it->second;
    uint32_t* p = (uint32_t*)alloc((1ull << (log - 1)) * 4);
    launch_circle_inv_y(*this, log, p);
    circle_inv_y[log] = p;
    return p;
}

}  // namespace sr
