// Lifted Blake2s Merkle commitment on gfx950.
//
// Leaves (src/prover/vcs_lifted/leaves.zig:24-200; blake2_merkle.zig:193-215):
//   leaf[p] = Blake2s("leaf"+0^60 || LE32(col[lift(p)]) for columns sorted by
//   (log size asc, original index)), lift(p) = ((p >> (L-l+1)) << 1) + (p & 1).
// One leaf per lane; the 64-byte prefix is pre-absorbed into a seed state, then
// 16 column words per compression are read with coalesced loads (consecutive
// lanes read consecutive rows of the same column).
// Shared prefixes: with the columns sorted by size, the hash state after message
// block k depends only on the lifted row at level M_k, the largest log among the
// columns of blocks 0..k (lifts compose: lifting L -> M then M -> l equals
// lifting L -> l).  Blocks of equal M_k form a segment; each segment is one
// launch over its 2^M_k states that starts from the previous segment's states
// (lifted) and writes its own, so a block is compressed 2^M_k times instead of
// 2^L times.  The result is the same leaf hash.
//
// Parents (layers.zig:99-140; blake2_merkle.zig:112-154): node = Blake2s(
// "node"+0^60 || left || right), one compression from the node seed.  One
// launch builds up to 10 levels: each thread hashes four children from HBM
// into two parents and their parent in registers, then the block halves the
// level in LDS, writing every level; levels of at most 64 nodes use one quad of
// lanes per compression.  The last 16 or fewer levels are one tree-layer launch
// (kernels_fri.hip) whose workgroups reduce up to 2^10 hashes each and whose last
// workgroup reduces their roots.  A 2^21-leaf tree takes two launches instead of 21.
#include <algorithm>
#include <vector>

#include "chan.hpp"

namespace sr {

constexpr int MK_THREADS = 256;

__device__ __forceinline__ void store_hash(uint8_t* __restrict__ out, uint64_t i, const uint32_t h[8]) {
    uint4* o = (uint4*)(out + i * 32);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

// boff: byte offset of the proof's columns in a batch (0 for one tree).
__device__ __forceinline__ void leaf_block_load(const LeafColDev* __restrict__ cd, uint32_t p, uint32_t m[16],
                                                uint64_t boff) {
#pragma unroll
    for (int w = 0; w < 16; w++) {
        const LeafColDev d = cd[w];
        m[w] = ((gcu32*)((const uint8_t*)d.col + boff))[((p >> d.shift) << 1) + (p & 1)] & d.mask;
    }
}

// The same for a segment whose columns all have the segment's size (lift(p) = p): every
// word is the column pointer (uniform) plus the lane's byte offset p * 4 (rows < 2^30), and
// the padding mask applies only to a padded last block (MASKED).
template <bool MASKED>
__device__ __forceinline__ void leaf_block_load_ident(const LeafColDev* __restrict__ cd, uint32_t voff,
                                                      uint32_t m[16], uint64_t boff) {
#pragma unroll
    for (int w = 0; w < 16; w++) {
        const LeafColDev d = cd[w];
        const uint32_t x = *(gcu32*)((const SR_GLOBAL uint8_t*)d.col + boff + voff);
        m[w] = MASKED ? (x & d.mask) : x;
    }
}

// One state per lane over blocks [blk0, blk0 + nblk) of a segment at `level`: the
// lane starts from the seed (first segment) or from the previous segment's state
// at the lifted row, loads the 16 words of the next block while the current one is
// compressed, and stores the state (the leaf hash after the last block).  IDENT: every
// column of the segment has its size (no lift; the blocks alternate between two buffers).
template <bool IDENT>
__global__ __launch_bounds__(MK_THREADS) void merkle_leaves_kernel(const LeafColDev* __restrict__ cols,
                                                                   uint32_t blk0, uint32_t nblk, uint32_t n_cols,
                                                                   uint32_t level, const uint32_t* __restrict__ in,
                                                                   uint32_t in_shift, B2State seed,
                                                                   uint8_t* __restrict__ out, uint64_t bstride) {
    const uint64_t p64 = (uint64_t)blockIdx.x * MK_THREADS + threadIdx.x;
    if (p64 >= (1ull << level)) return;
    const uint32_t p = (uint32_t)p64;
    // tree blockIdx.y of a batch: its columns, states and hashes bstride bytes further on
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    out += boff;
    if (in) in = (const uint32_t*)((const uint8_t*)in + boff);
    uint32_t h[8];
    if (in) {
        const SR_GLOBAL u32x4* s4 = (const SR_GLOBAL u32x4*)((gcu32*)in + 8ull * (((p >> in_shift) << 1) + (p & 1)));
        const u32x4 a = s4[0], b = s4[1];
        h[0] = a.x, h[1] = a.y, h[2] = a.z, h[3] = a.w, h[4] = b.x, h[5] = b.y, h[6] = b.z, h[7] = b.w;
    } else {
#pragma unroll
        for (int i = 0; i < 8; i++) h[i] = seed.h[i];
    }
    const uint32_t nblocks_total = (n_cols + 15) >> 4;
    const uint32_t blk_end = blk0 + nblk;
    if constexpr (IDENT) {
        // blocks of 16 real columns, then at most one padded block (the last of the leaf)
        const uint32_t full_end = ((n_cols & 15) && blk_end == nblocks_total) ? blk_end - 1 : blk_end;
        const uint32_t voff = p << 2;
        uint32_t a[16], b[16];
        uint32_t blk = blk0;
        if (blk < full_end) leaf_block_load_ident<false>(cols + 16 * blk, voff, a, boff);
        for (; blk + 1 < full_end; blk += 2) {
            leaf_block_load_ident<false>(cols + 16 * (blk + 1), voff, b, boff);
            b2_compress(h, a, 64 + 64 * (blk + 1), blk + 1 == nblocks_total);
            if (blk + 2 < full_end) leaf_block_load_ident<false>(cols + 16 * (blk + 2), voff, a, boff);
            b2_compress(h, b, 64 + 64 * (blk + 2), blk + 2 == nblocks_total);
        }
        if (blk < full_end) {
            b2_compress(h, a, 64 + 64 * (blk + 1), blk + 1 == nblocks_total);
            blk++;
        }
        if (blk < blk_end) {
            leaf_block_load_ident<true>(cols + 16 * blk, voff, a, boff);
            b2_compress(h, a, 64 + 4 * n_cols, true);
        }
    } else {
        uint32_t cur[16];
        leaf_block_load(cols + 16 * blk0, p, cur, boff);
        for (uint32_t blk = blk0; blk < blk_end; blk++) {
            uint32_t nxt[16];
            const bool more = blk + 1 < blk_end;
            if (more) leaf_block_load(cols + 16 * (blk + 1), p, nxt, boff);
            const bool last = blk + 1 == nblocks_total;
            const uint32_t t = 64 + (last ? 4 * n_cols : 64 * (blk + 1));
            b2_compress(h, cur, t, last);
            if (more) {
#pragma unroll
                for (int w = 0; w < 16; w++) cur[w] = nxt[w];
            }
        }
    }
    store_hash(out, p, h);
}

// The same with one leaf per quad of lanes (one compression spread over four lanes): for few
// leaves with many message blocks (a wide trace on a small domain) the per-leaf chain of
// compressions, not the lane count, sets the time.  Every lane loads the whole block.
__global__ __launch_bounds__(MK_THREADS) void merkle_leaves_quad_kernel(const LeafColDev* __restrict__ cols,
                                                                        uint32_t blk0, uint32_t nblk, uint32_t n_cols,
                                                                        uint32_t level, const uint32_t* __restrict__ in,
                                                                        uint32_t in_shift, B2State seed,
                                                                        uint8_t* __restrict__ out, uint64_t bstride) {
    const uint64_t t64 = (uint64_t)blockIdx.x * MK_THREADS + threadIdx.x;
    if (t64 >= (4ull << level)) return;  // whole quads: the thread count is a multiple of 4
    const uint32_t p = (uint32_t)(t64 >> 2), q = (uint32_t)(t64 & 3);
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    out += boff;
    if (in) in = (const uint32_t*)((const uint8_t*)in + boff);
    const QuadLane L(q);
    uint32_t hlo, hhi;
    if (in) {
        const uint32_t* s = in + 8ull * (((p >> in_shift) << 1) + (p & 1));
        hlo = s[q];
        hhi = s[4 + q];
    } else {
        hlo = L.sel(seed.h[0], seed.h[1], seed.h[2], seed.h[3]);
        hhi = L.sel(seed.h[4], seed.h[5], seed.h[6], seed.h[7]);
    }
    const uint32_t nblocks_total = (n_cols + 15) >> 4;
    const uint32_t blk_end = blk0 + nblk;
    uint32_t cur[16];
    leaf_block_load(cols + 16 * blk0, p, cur, boff);
    for (uint32_t blk = blk0; blk < blk_end; blk++) {
        uint32_t nxt[16];
        const bool more = blk + 1 < blk_end;
        if (more) leaf_block_load(cols + 16 * (blk + 1), p, nxt, boff);
        const bool last = blk + 1 == nblocks_total;
        const uint32_t t = 64 + (last ? 4 * n_cols : 64 * (blk + 1));
        b2_compress_quad(L, hlo, hhi, cur, t, last);
        if (more) {
#pragma unroll
            for (int w = 0; w < 16; w++) cur[w] = nxt[w];
        }
    }
    uint32_t* o = (uint32_t*)(out + (uint64_t)p * 32);
    o[q] = hlo;
    o[4 + q] = hhi;
}

// Leaves of a QM31 column in AoS form: 4 coordinate words per leaf (FRI layers).
__global__ __launch_bounds__(MK_THREADS) void merkle_leaves_qm31_kernel(const uint4* __restrict__ vals,
                                                                        uint32_t log_size, B2State seed,
                                                                        uint8_t* __restrict__ leaves) {
    uint64_t p = (uint64_t)blockIdx.x * MK_THREADS + threadIdx.x;
    if (p >= (1ull << log_size)) return;
    uint4 v = vals[p];
    uint32_t m[16] = {v.x, v.y, v.z, v.w, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t h[8];
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = seed.h[i];
    b2_compress(h, m, 80, true);
    store_hash(leaves, p, h);
}

// Builds `levels` (2..10) parent levels below layer k.  Block b covers the
// 2^levels * ... children [b*4T, (b+1)*4T) of layer k for T = 2^(levels-2)
// threads: every thread hashes two parents of level k-1 from four children
// (two independent compressions) and their parent of level k-2 in registers,
// so the first two levels keep every lane busy; the remaining levels halve in
// LDS.  levels == 1 uses one thread per parent.
constexpr int MP_THREADS = 256;
__device__ __forceinline__ void node_hash(const B2State& seed, const uint32_t* l, const uint32_t* r, uint32_t h[8]) {
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = l[i];
        m[8 + i] = r[i];
    }
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = seed.h[i];
    b2_compress(h, m, 128, true);
}

// Levels k-1 .. k-levels of a tree: block b reduces its 2^levels hashes of level k.
__global__ __launch_bounds__(MP_THREADS) void merkle_parents_kernel(uint8_t* __restrict__ tree, uint32_t k,
                                                                    uint32_t levels, B2State seed, uint64_t bstride) {
    __shared__ uint32_t sh[MP_THREADS * 8];
    const uint32_t t = threadIdx.x;
    tree += (uint64_t)blockIdx.y * bstride;  // tree blockIdx.y of a batch
    const uint8_t* child = tree + (((uint64_t)1 << k) - 1) * 32;
    if (levels == 1) {
        uint64_t i = (uint64_t)blockIdx.x * MP_THREADS + t;
        if (i >= (1ull << (k - 1))) return;
        const uint4* c4 = (const uint4*)(child + i * 64);
        uint4 a = c4[0], b = c4[1], c = c4[2], d = c4[3];
        uint32_t l[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w}, r[8] = {c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
        uint32_t h[8];
        node_hash(seed, l, r, h);
        store_hash(tree + (((uint64_t)1 << (k - 1)) - 1) * 32, i, h);
        return;
    }
    const uint32_t T = 1u << (levels - 2);  // threads with work (<= 256)
    uint64_t g = (uint64_t)blockIdx.x * T + t;  // index at level k-2
    uint32_t h2[8];
    if (t < T) {
        const uint4* c4 = (const uint4*)(child + g * 128);
        uint32_t c[32];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            uint4 v = c4[q];
            c[4 * q] = v.x;
            c[4 * q + 1] = v.y;
            c[4 * q + 2] = v.z;
            c[4 * q + 3] = v.w;
        }
        uint32_t p0[8], p1[8];
        node_hash(seed, c, c + 8, p0);
        node_hash(seed, c + 16, c + 24, p1);
        uint8_t* l1 = tree + (((uint64_t)1 << (k - 1)) - 1) * 32;
        store_hash(l1, 2 * g, p0);
        store_hash(l1, 2 * g + 1, p1);
        node_hash(seed, p0, p1, h2);
        store_hash(tree + (((uint64_t)1 << (k - 2)) - 1) * 32, g, h2);
#pragma unroll
        for (int i = 0; i < 8; i++) sh[t * 8 + i] = h2[i];
    }
    uint32_t count = T, level = k - 2;
    uint64_t first = (uint64_t)blockIdx.x * T;
    const uint32_t q = t & 3, j = t >> 2;
    const QuadLane L(q);
    for (uint32_t lv = 2; lv < levels; lv++) {
        __syncthreads();
        count >>= 1;
        level--;
        first >>= 1;
        uint8_t* out = tree + (((uint64_t)1 << level) - 1) * 32;
        if (count > blockDim.x / 4) {
            uint32_t h[8];
            if (t < count) node_hash(seed, &sh[t * 16], &sh[t * 16 + 8], h);
            __syncthreads();
            if (t < count) {
#pragma unroll
                for (int i = 0; i < 8; i++) sh[t * 8 + i] = h[i];
                store_hash(out, first + t, h);
            }
        } else {
            // short levels: one compression per quad
            uint32_t hlo = 0, hhi = 0;
            const bool act = j < count;
            if (act) {
                uint32_t m[16];
#pragma unroll
                for (int i = 0; i < 16; i++) m[i] = sh[j * 16 + i];
                hlo = L.sel(seed.h[0], seed.h[1], seed.h[2], seed.h[3]);
                hhi = L.sel(seed.h[4], seed.h[5], seed.h[6], seed.h[7]);
                b2_compress_quad(L, hlo, hhi, m, 128, true);
            }
            __syncthreads();
            if (act) {
                sh[j * 8 + q] = hlo;
                sh[j * 8 + 4 + q] = hhi;
                uint32_t* o = (uint32_t*)(out + (first + j) * 32);
                o[q] = hlo;
                o[4 + q] = hhi;
            }
        }
    }
}

// Single parent layer from a separate children array (commit_on_layer shape).
__global__ __launch_bounds__(MK_THREADS) void merkle_parent_layer_kernel(const uint8_t* __restrict__ children,
                                                                         uint32_t log_size, B2State seed,
                                                                         uint8_t* __restrict__ parents) {
    uint64_t i = (uint64_t)blockIdx.x * MK_THREADS + threadIdx.x;
    if (i >= (1ull << log_size)) return;
    const uint4* c4 = (const uint4*)(children + i * 64);
    uint4 a = c4[0], b = c4[1], c = c4[2], d = c4[3];
    uint32_t m[16] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
    uint32_t h[8];
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = seed.h[k];
    b2_compress(h, m, 128, true);
    store_hash(parents, i, h);
}

void launch_merkle_parent_layer(Ctx& c, const uint8_t* d_children, uint32_t log_size, uint8_t* d_parents) {
    uint64_t n = 1ull << log_size;
    hipLaunchKernelGGL(merkle_parent_layer_kernel, dim3((uint32_t)((n + MK_THREADS - 1) / MK_THREADS)),
                       dim3(MK_THREADS), 0, c.stream, d_children, log_size, c.node_seed, d_parents);
    SR_HIP(hipGetLastError());
}

LeafPlan merkle_leaf_plan(const uint32_t* const* h_cols, const uint32_t* h_logs, uint32_t n_cols) {
    SR_CHECK(n_cols > 0, "merkle leaves: no columns");
    const uint32_t nblocks = (n_cols + 15) / 16;
    LeafPlan plan;
    plan.tab.resize(16ull * nblocks);
    // block k's level: the running maximum of the column logs up to the end of block k
    uint32_t level = 0;
    for (uint32_t k = 0; k < nblocks; k++) {
        for (uint32_t i = 16 * k; i < std::min(n_cols, 16 * k + 16); i++) {
            SR_CHECK(h_logs[i] >= 1 && h_logs[i] <= 30, "merkle leaves: column log size");
            SR_CHECK(i == 0 || h_logs[i - 1] <= h_logs[i], "merkle leaves: columns must be sorted by log size");
            level = std::max(level, h_logs[i]);
        }
        if (plan.segs.empty() || plan.segs.back().level != level) plan.segs.push_back(LeafSeg{k, 0, level, 0});
        LeafSeg& sg = plan.segs.back();
        sg.nblk++;
        for (uint32_t w = 0; w < 16; w++) {
            const uint32_t i = 16 * k + w;
            // padding (only in the last block, of the final segment) reads the last column, whose
            // log is the final level: in range for both load forms, then masked to zero
            const bool real = i < n_cols;
            const uint32_t lg = real ? h_logs[i] : h_logs[n_cols - 1];
            LeafColDev& d = plan.tab[i];
            d.col = real ? h_cols[i] : h_cols[n_cols - 1];
            d.shift = level - lg + 1;
            d.mask = real ? 0xffffffffu : 0u;
            if (real) sg.col_words += 1ull << lg;
            if (real && d.shift != 1) sg.ident = false;
        }
    }
    return plan;
}

void launch_merkle_leaves_dev(Ctx& c, const LeafColDev* d_tab, const std::vector<LeafSeg>& segs, uint32_t n_cols,
                              DeviceTree& t, uint32_t batch, uint64_t bstride) {
    const uint32_t L = t.max_log;
    SR_CHECK(batch >= 1 && batch <= 65535, "merkle leaves: batch size");
    const size_t S = segs.size();
    SR_CHECK(S >= 1 && segs.back().level == L && segs.back().blk0 + segs.back().nblk == (n_cols + 15) / 16,
             "merkle leaves: segment plan does not match the tree");
    const uint32_t* prev = nullptr;
    uint32_t prev_level = 0;
    for (size_t s = 0; s < S; s++) {
        const LeafSeg& sg = segs[s];
        SR_CHECK(s == 0 || sg.level > prev_level, "merkle leaves: segment levels must increase");
        const bool final = s + 1 == S;
        // intermediate states: layers L-1 and L-2 alternately, counted back from the final
        // segment (level M_s <= L - (S-1-s) fits layer L-1 or L-2 by parity); the parent
        // launches overwrite both afterwards
        uint8_t* out = final ? t.layer(L) : t.layer(((S - 1 - s) & 1) ? L - 1 : L - 2);
        const uint64_t n = 1ull << sg.level;
        double bytes = batch * (4.0 * (double)sg.col_words + 32.0 * (double)n +
                                (prev ? 32.0 * (double)(1ull << prev_level) : 0.0));
        KTimer kt(c,
                  final ? "merkle_leaves L=" + std::to_string(L) + " cols=" + std::to_string(n_cols) +
                              (S > 1 ? " blocks=" + std::to_string(sg.nblk) : "")
                        : "merkle_leaves_prefix L=" + std::to_string(sg.level) + " blocks=" + std::to_string(sg.nblk),
                  bytes);
        if (n <= (1u << 14) && sg.nblk >= 8) {
            hipLaunchKernelGGL(merkle_leaves_quad_kernel,
                               dim3((uint32_t)((4 * n + MK_THREADS - 1) / MK_THREADS), batch), dim3(MK_THREADS), 0,
                               c.stream, d_tab, sg.blk0, sg.nblk, n_cols, sg.level, prev,
                               prev ? sg.level - prev_level + 1 : 0u, c.leaf_seed, out, bstride);
        } else {
            const dim3 grid((uint32_t)((n + MK_THREADS - 1) / MK_THREADS), batch);
            if (sg.ident)
                hipLaunchKernelGGL(merkle_leaves_kernel<true>, grid, dim3(MK_THREADS), 0, c.stream, d_tab, sg.blk0,
                                   sg.nblk, n_cols, sg.level, prev, prev ? sg.level - prev_level + 1 : 0u,
                                   c.leaf_seed, out, bstride);
            else
                hipLaunchKernelGGL(merkle_leaves_kernel<false>, grid, dim3(MK_THREADS), 0, c.stream, d_tab, sg.blk0,
                                   sg.nblk, n_cols, sg.level, prev, prev ? sg.level - prev_level + 1 : 0u,
                                   c.leaf_seed, out, bstride);
        }
        SR_HIP(hipGetLastError());
        prev = (const uint32_t*)out;
        prev_level = sg.level;
    }
}

void launch_merkle_leaves(Ctx& c, const uint32_t* const* h_cols, const uint32_t* h_logs, uint32_t n_cols,
                          DeviceTree& t) {
    LeafPlan plan = merkle_leaf_plan(h_cols, h_logs, n_cols);
    DBuf d_tab = upload(c, plan.tab.data(), plan.tab.size());
    launch_merkle_leaves_dev(c, d_tab.as<LeafColDev>(), plan.segs, n_cols, t);
}

void launch_merkle_leaves_qm31(Ctx& c, const uint32_t* d_qm31, uint32_t log_size, uint8_t* d_leaves) {
    uint64_t n = 1ull << log_size;
    KTimer kt(c, "merkle_leaves_qm31", 48.0 * (double)n);
    hipLaunchKernelGGL(merkle_leaves_qm31_kernel, dim3((uint32_t)((n + MK_THREADS - 1) / MK_THREADS)),
                       dim3(MK_THREADS), 0, c.stream, (const uint4*)d_qm31, log_size, c.leaf_seed, d_leaves);
    SR_HIP(hipGetLastError());
}

// Wide parent levels (throughput-bound) with merkle_parents_kernel while more than
// TREE_LAYER_MAX_LOG levels remain, then the top levels and the channel step in one
// tree-layer launch (latency-bound: one quad of lanes per compression on the narrow levels).
void launch_merkle_parents(Ctx& c, DeviceTree& t, DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch,
                           uint64_t bstride, uint32_t* d_cnt) {
    uint32_t k = t.max_log;
    SR_CHECK(batch >= 1 && batch <= 65535, "merkle parents: batch size");
    if (k == 0) {
        SR_CHECK(batch == 1, "merkle parents: a batch of one-leaf trees");
        if (d_ch) launch_chan_mix_root_draw(c, d_ch, t.layers, d_root_out);
        else if (d_root_out) SR_HIP(hipMemcpyAsync(d_root_out, t.layers, 32, hipMemcpyDeviceToDevice, c.stream));
        return;
    }
    while (k > TREE_LAYER_MAX_LOG) {
        // at least 8 levels stay for the top launch (up to 2^8 hashes reduce in one workgroup)
        const uint32_t levels = std::min(c.merkle_levels, k - 8);
        const uint32_t blocks = levels == 1 ? (uint32_t)(((1ull << (k - 1)) + MP_THREADS - 1) / MP_THREADS)
                                            : (1u << (k - levels));
        // levels >= 2: 2^(levels-2) threads carry work; at least one wave per block
        const uint32_t threads = levels == 1 ? MP_THREADS : std::max(64u, 1u << (levels - 2));
        KTimer kt(c, "merkle_parents L=" + std::to_string(k) + " levels=" + std::to_string(levels),
                  batch * 32.0 * (double)((1ull << k) + (1ull << k) - (1ull << (k - levels))));
        hipLaunchKernelGGL(merkle_parents_kernel, dim3(blocks, batch), dim3(threads), 0, c.stream, t.layers, k, levels,
                           c.node_seed, bstride);
        SR_HIP(hipGetLastError());
        k -= levels;
    }
    launch_tree_top(c, t, k, d_ch, d_root_out, batch, bstride, d_cnt);
}

static DeviceTree* new_tree(Ctx& c, uint32_t max_log) {
    SR_CHECK(max_log <= 30, "merkle: log size too large");
    DeviceTree* t = new DeviceTree();
    t->max_log = max_log;
    t->layers = (uint8_t*)c.alloc(((2ull << max_log) - 1) * 32);
    return t;
}

DeviceTree* merkle_commit(Ctx& c, const std::vector<const uint32_t*>& cols, const std::vector<uint32_t>& logs) {
    SR_CHECK(cols.size() == logs.size(), "merkle: column/log count mismatch");
    std::vector<uint32_t> order(cols.size());
    for (size_t i = 0; i < order.size(); i++) {
        SR_CHECK(logs[i] >= 1 && logs[i] <= 30, "merkle: column length must be a power of two >= 2");
        order[i] = (uint32_t)i;
    }
    // sortByLogSizeAsc, stable on the original index (columns.zig:12-46).
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return logs[a] < logs[b]; });
    uint32_t max_log = order.empty() ? 0 : logs[order.back()];
    DeviceTree* t = new_tree(c, max_log);
    if (cols.empty()) {
        // No columns: a single leaf Blake2s(prefix) (leaves.zig:30-34).
        uint8_t prefix[64] = {'l', 'e', 'a', 'f'};
        uint8_t h[32];
        b2_hash(prefix, 64, h);
        SR_HIP(hipMemcpyAsync(t->layers, c.stage(h, 32), 32, hipMemcpyHostToDevice, c.stream));
        memcpy(t->root, h, 32);
        return t;
    }
    std::vector<const uint32_t*> sc(cols.size());
    std::vector<uint32_t> sl(cols.size());
    for (size_t i = 0; i < order.size(); i++) {
        sc[i] = cols[order[i]];
        sl[i] = logs[order[i]];
    }
    launch_merkle_leaves(c, sc.data(), sl.data(), (uint32_t)sc.size(), *t);
    launch_merkle_parents(c, *t);
    return t;
}

DeviceTree* merkle_commit_qm31(Ctx& c, const uint32_t* d_qm31, uint32_t log_size) {
    DeviceTree* t = new_tree(c, log_size);
    launch_merkle_leaves_qm31(c, d_qm31, log_size, t->layer(log_size));
    launch_merkle_parents(c, *t);
    return t;
}

void tree_read_root(Ctx& c, DeviceTree& t) {
    uint8_t* h = c.pinned(32);
    SR_HIP(hipMemcpyAsync(h, t.layers, 32, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    memcpy(t.root, h, 32);
}

void tree_free(Ctx& c, DeviceTree* t) {
    if (!t) return;
    if (t->layers) c.free(t->layers);
    delete t;
}

}  // namespace sr
