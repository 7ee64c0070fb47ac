// FRI folds on gfx950 (src/core/fri/folding.zig:316-380,558-602).
//
// fold_line:   dst[i] = (f0 + f1) + alpha * (f0 - f1) * x_i^-1, (f0, f1) = src[2i], src[2i+1],
//              x_i = LineDomain(halfOdds(log)).at(bitrev(i)).x.  x_i^-1 is read from the
//              inverse twiddle tower: its layer for halfOdds(log) holds exactly these values.
// fold_circle: dst[i] = dst[i] * alpha^2 + (f0 + f1) + alpha * (f0 - f1) * y_i^-1 with
//              y_i = halfOdds(log-1).at(bitrev(i)).y (resident table per log size).
// Both optionally hash the folded value as the next layer's Merkle leaf
// (prefix || 4 LE words, one compression), fusing fold and leaf hashing.
//
// tree_layer: one FRI layer commitment per launch (prover/fri.zig:377-621): the fold into
// the layer's line, its leaf hashes, every parent level (layers.zig:99-140) and the channel
// step on the root; the top levels are reduced by the last workgroup to finish, so the next
// layer's launch finds its alpha in the device channel.  The same kernel reduces the top
// levels of any tree from hashes already in HBM.
// fri_tail: the layers below FRI_TAIL_MAX_LOG + 1 in one workgroup with the line and the
// hash level resident in LDS.
#include "chan.hpp"

namespace sr {

__device__ __forceinline__ void leaf_qm31(uint8_t* __restrict__ leaves, uint64_t i, const B2State& seed, qm31 v) {
    uint32_t m[16] = {v.c0.a, v.c0.b, v.c1.a, v.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t h[8];
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = seed.h[k];
    b2_compress(h, m, 80, true);
    uint4* o = (uint4*)(leaves + i * 32);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

__global__ __launch_bounds__(256) void fold_line_kernel(const uint4* __restrict__ src, uint32_t log_src,
                                                        const uint32_t* __restrict__ inv_x, qm31 alpha,
                                                        const DevChannel* __restrict__ dch,
                                                        uint4* __restrict__ dst, uint8_t* __restrict__ leaves,
                                                        B2State seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << (log_src - 1))) return;
    if (dch) alpha = qload(dch->alpha);
    uint4 a = src[2 * i], b = src[2 * i + 1];
    qm31 f0 = {{a.x, a.y}, {a.z, a.w}}, f1 = {{b.x, b.y}, {b.z, b.w}};
    qm31 r = qadd(qadd(f0, f1), qmul(qmulm(qsub(f0, f1), inv_x[i]), alpha));
    dst[i] = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
    if (leaves) leaf_qm31(leaves, i, seed, r);
}

__global__ __launch_bounds__(256) void fold_circle_kernel(const uint32_t* __restrict__ c0,
                                                          const uint32_t* __restrict__ c1,
                                                          const uint32_t* __restrict__ c2,
                                                          const uint32_t* __restrict__ c3, uint32_t log_src,
                                                          const uint32_t* __restrict__ inv_y, qm31 alpha, qm31 alpha2,
                                                          const DevChannel* __restrict__ dch, bool accumulate,
                                                          uint4* __restrict__ dst, uint8_t* __restrict__ leaves,
                                                          B2State seed, uint64_t bstride) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << (log_src - 1))) return;
    if (blockIdx.y) {  // proof blockIdx.y of a batch: its buffers bstride bytes further on
        const uint64_t boff = (uint64_t)blockIdx.y * bstride;
        c0 = (const uint32_t*)((const uint8_t*)c0 + boff);
        c1 = (const uint32_t*)((const uint8_t*)c1 + boff);
        c2 = (const uint32_t*)((const uint8_t*)c2 + boff);
        c3 = (const uint32_t*)((const uint8_t*)c3 + boff);
        if (dch) dch = (const DevChannel*)((const uint8_t*)dch + boff);
        dst = (uint4*)((uint8_t*)dst + boff);
        if (leaves) leaves += boff;
    }
    if (dch) {
        alpha = qload(dch->alpha);
        alpha2 = qload(dch->alpha2);
    }
    uint2 w0 = ((const uint2*)c0)[i], w1 = ((const uint2*)c1)[i], w2 = ((const uint2*)c2)[i],
          w3 = ((const uint2*)c3)[i];
    qm31 f0 = {{w0.x, w1.x}, {w2.x, w3.x}}, f1 = {{w0.y, w1.y}, {w2.y, w3.y}};
    qm31 r = qadd(qadd(f0, f1), qmul(qmulm(qsub(f0, f1), inv_y[i]), alpha));
    if (accumulate) {
        uint4 d = dst[i];
        r = qadd(qmul(qm31{{d.x, d.y}, {d.z, d.w}}, alpha2), r);
    }
    dst[i] = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
    if (leaves) leaf_qm31(leaves, i, seed, r);
}

// ---------------------------------------------------------------------------
// Tree levels inside one workgroup.  A level of c <= 128 parents uses one quad of lanes per
// compression (blake2s_quad.hpp: about half the latency of a one-lane compression while a
// SIMD runs one wave); wider levels use one lane per compression, which does the same
// work in fewer instructions once the SIMDs are full.  Shared by the layer and tail kernels.

constexpr uint32_t TL_THREADS = 256;    // tree-layer workgroup: one wave per SIMD
constexpr uint32_t TAIL_THREADS = 1024;  // tail workgroup: a whole line of 2^10 values
constexpr uint32_t TR_QUAD_MAX = 128;    // widest level hashed by quads

__device__ __forceinline__ void tr_load16(const uint32_t* in, uint32_t m[16]) {
    const uint4* m4 = (const uint4*)in;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint4 v = m4[i];
        m[4 * i] = v.x, m[4 * i + 1] = v.y, m[4 * i + 2] = v.z, m[4 * i + 3] = v.w;
    }
}

__device__ __forceinline__ void tr_put_quad(uint32_t* lds, uint8_t* __restrict__ g, uint32_t q, uint32_t hlo,
                                            uint32_t hhi) {
    lds[q] = hlo;
    lds[4 + q] = hhi;
    uint32_t* o = (uint32_t*)g;
    o[q] = hlo;
    o[4 + q] = hhi;
}

__device__ __forceinline__ void tr_put_lane(uint32_t* lds, uint8_t* __restrict__ g, const uint32_t h[8]) {
    const uint4 x = make_uint4(h[0], h[1], h[2], h[3]), y = make_uint4(h[4], h[5], h[6], h[7]);
    *(uint4*)lds = x;
    *(uint4*)(lds + 4) = y;
    uint4* o = (uint4*)g;
    o[0] = x;
    o[1] = y;
}

// Parents p < c of the 2c child hashes in `in` (LDS, 8 words per hash), written to `out`
// (LDS) and to the tree level `gl` at hashes first + p (layers.zig:99-140:
// H(node_seed; left || right)).  The level of one parent (c == 1) runs on quad 0, which is
// left holding it as (rlo, rhi) = words (q, 4 + q).  Ends with a workgroup barrier.
template <uint32_t T>
__device__ __forceinline__ void tr_level(const uint32_t* in, uint32_t* out, uint32_t c, uint8_t* __restrict__ gl,
                                         uint64_t first, const B2State& seed, uint32_t t, uint32_t& rlo,
                                         uint32_t& rhi) {
    if (c <= TR_QUAD_MAX) {
        const uint32_t q = t & 3;
        const QuadLane L(q);
        for (uint32_t p = t >> 2; p < c; p += T / 4) {
            uint32_t m[16];
            tr_load16(in + 16 * p, m);
            uint32_t hlo = L.sel(seed.h[0], seed.h[1], seed.h[2], seed.h[3]);
            uint32_t hhi = L.sel(seed.h[4], seed.h[5], seed.h[6], seed.h[7]);
            b2_compress_quad(L, hlo, hhi, m, 128, true);
            tr_put_quad(out + 8 * p, gl + (first + p) * 32, q, hlo, hhi);
            if (p == 0) {
                rlo = hlo;
                rhi = hhi;
            }
        }
    } else {
        for (uint32_t p = t; p < c; p += T) {
            uint32_t m[16], h[8];
            tr_load16(in + 16 * p, m);
#pragma unroll
            for (int k = 0; k < 8; k++) h[k] = seed.h[k];
            b2_compress(h, m, 128, true);
            tr_put_lane(out + 8 * p, gl + (first + p) * 32, h);
        }
    }
    __syncthreads();
}

// Leaf hash of one QM31 value (leaf_qm31 above) by a quad, the value in every lane.
__device__ __forceinline__ void tr_leaf_quad(const QuadLane& L, const B2State& seed, qm31 v, uint32_t& hlo,
                                             uint32_t& hhi) {
    const uint32_t m[16] = {v.c0.a, v.c0.b, v.c1.a, v.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    hlo = L.sel(seed.h[0], seed.h[1], seed.h[2], seed.h[3]);
    hhi = L.sel(seed.h[4], seed.h[5], seed.h[6], seed.h[7]);
    b2_compress_quad(L, hlo, hhi, m, 80, true);
}

__device__ __forceinline__ void tr_leaf_lane(const B2State& seed, qm31 v, uint32_t h[8]) {
    const uint32_t m[16] = {v.c0.a, v.c0.b, v.c1.a, v.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = seed.h[k];
    b2_compress(h, m, 80, true);
}

__device__ __forceinline__ qm31 tr_fold(uint4 a, uint4 b, qm31 alpha, uint32_t inv) {
    const qm31 f0 = {{a.x, a.y}, {a.z, a.w}}, f1 = {{b.x, b.y}, {b.z, b.w}};
    return qadd(qadd(f0, f1), qmul(qmulm(qsub(f0, f1), inv), alpha));
}

// Fold input i of a layer kernel: line (fold_line_kernel) or circle (fold_circle_kernel,
// no accumulation) source.
template <int SRC>
__device__ __forceinline__ qm31 tr_fold_at(const FriLayerArgs& a, uint64_t i, qm31 alpha) {
    if constexpr (SRC == FRI_SRC_LINE) {
        return tr_fold(a.src_line[2 * i], a.src_line[2 * i + 1], alpha, a.inv[i]);
    } else {
        const uint2 w0 = ((const uint2*)a.coord[0])[i], w1 = ((const uint2*)a.coord[1])[i],
                    w2 = ((const uint2*)a.coord[2])[i], w3 = ((const uint2*)a.coord[3])[i];
        return tr_fold(make_uint4(w0.x, w1.x, w2.x, w3.x), make_uint4(w0.y, w1.y, w2.y, w3.y), alpha, a.inv[i]);
    }
}

// Channel step on a root held by quad 0 as (rlo, rhi), the channel digest `dg` in registers
// of every lane of the quad (MerkleChannel.mixRoot then drawSecureFelt, as
// chan_mix_root_draw_quad): updates dg and the DevChannel, copies the root to root_out.
__device__ __forceinline__ qm31 tr_chan_step(const QuadLane& L, uint32_t q, uint32_t dg[8], uint32_t rlo, uint32_t rhi,
                                             DevChannel* __restrict__ ch, uint32_t* __restrict__ root_out) {
    uint32_t pay[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    quad_gather8(rlo, rhi, pay);
    if (root_out) {
        root_out[q] = rlo;
        root_out[4 + q] = rhi;
    }
    qchan_mix(L, dg, pay, 8);
    uint32_t nd = 0, st = 0;
    const qm31 a = qchan_draw(L, dg, nd, st);
    // lane q stores digest words q and 4 + q (the digest is in every lane)
    ch->digest[q] = L.sel(dg[0], dg[1], dg[2], dg[3]);
    ch->digest[4 + q] = L.sel(dg[4], dg[5], dg[6], dg[7]);
    if (q == 0) {
        ch->n_draws = nd;
        qstore(a, ch->alpha);
        qstore(qsqr(a), ch->alpha2);
        if (st) ch->status = 1;
    }
    return a;
}

// One tree (or its top levels) in one launch.  Workgroup b owns the 2^s hashes
// [b 2^s, (b+1) 2^s) of level k0: it produces them (SRC = FRI_SRC_LINE / FRI_SRC_CIRCLE:
// fold the previous layer with the alpha in `ch`, store the folded value and hash it as a
// leaf) or loads them (FRI_SRC_HASHES), then reduces them to its subtree root at level
// k0 - s.  The last workgroup to finish (split completion counters; the subtree roots are
// exchanged through device-coherent stores and loads, so no workgroup writes back its L2)
// reduces the 2^(k0-s) subtree roots to the root and, with `ch`, runs the channel step on
// it.  The counters are left at zero.
template <int SRC>
__global__ __launch_bounds__(TL_THREADS) void tree_layer_kernel(FriLayerArgs a0) {
    FriLayerArgs a = a0;
    if (blockIdx.y) {
        // tree blockIdx.y of a batch: its line, tree, channel and root copy bstride bytes further
        // on, its counters and root exchange TREE_CNT_WORDS words further on
        const uint64_t boff = (uint64_t)blockIdx.y * a.bstride;
        if (a.src_line) a.src_line = (const uint4*)((const uint8_t*)a.src_line + boff);
        for (int k = 0; k < 4; k++)
            if (a.coord[k]) a.coord[k] = (const uint32_t*)((const uint8_t*)a.coord[k] + boff);
        if (a.dst_line) a.dst_line = (uint4*)((uint8_t*)a.dst_line + boff);
        a.tree += boff;
        if (a.ch) a.ch = (DevChannel*)((uint8_t*)a.ch + boff);
        if (a.root_out) a.root_out = (uint32_t*)((uint8_t*)a.root_out + boff);
        a.cnt += (uint64_t)blockIdx.y * TREE_CNT_WORDS;
        a.xr += (uint64_t)blockIdx.y * TREE_CNT_WORDS;
    }
    __shared__ __align__(16) uint32_t hA[1024 * 8];
    __shared__ __align__(16) uint32_t hB[512 * 8];
    __shared__ uint32_t last;
    const uint32_t t = threadIdx.x, q = t & 3;
    const QuadLane L(q);
    const uint32_t n1 = 1u << a.s;
    const uint64_t base = (uint64_t)blockIdx.x * n1;
    uint8_t* lev_k0 = a.tree + (((uint64_t)1 << a.k0) - 1) * 32;
    if constexpr (SRC == FRI_SRC_HASHES) {
        for (uint32_t p = t; p < n1; p += TL_THREADS) {
            const uint4* s4 = (const uint4*)(lev_k0 + (base + p) * 32);
            const uint4 x = s4[0], y = s4[1];
            *(uint4*)&hA[8 * p] = x;
            *(uint4*)&hA[8 * p + 4] = y;
        }
    } else {
        const qm31 alpha = qload(a.ch->alpha);
        if (n1 <= TR_QUAD_MAX) {
            for (uint32_t p = t >> 2; p < n1; p += TL_THREADS / 4) {
                const qm31 r = tr_fold_at<SRC>(a, base + p, alpha);
                if (q == 0) a.dst_line[base + p] = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
                uint32_t hlo, hhi;
                tr_leaf_quad(L, a.leaf_seed, r, hlo, hhi);
                tr_put_quad(hA + 8 * p, lev_k0 + (base + p) * 32, q, hlo, hhi);
            }
        } else {
            for (uint32_t p = t; p < n1; p += TL_THREADS) {
                const qm31 r = tr_fold_at<SRC>(a, base + p, alpha);
                a.dst_line[base + p] = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
                uint32_t h[8];
                tr_leaf_lane(a.leaf_seed, r, h);
                tr_put_lane(hA + 8 * p, lev_k0 + (base + p) * 32, h);
            }
        }
    }
    __syncthreads();
    uint32_t rlo = 0, rhi = 0;
    uint32_t *in = hA, *out = hB;
    uint32_t lev = a.k0;
    for (uint32_t c = n1 >> 1; c >= 1; c >>= 1) {
        lev--;
        tr_level<TL_THREADS>(in, out, c, a.tree + (((uint64_t)1 << lev) - 1) * 32, (uint64_t)blockIdx.x * c,
                             a.node_seed, t, rlo, rhi);
        uint32_t* x = in;
        in = out;
        out = x;
    }
    const uint32_t nb = gridDim.x;
    uint32_t dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (nb > 1) {
        // Hand-off without an L2 writeback: lanes 0..7 publish the subtree root (in LDS since
        // the barrier that ended the last level) with device-coherent stores (agent-scope
        // atomic stores, global_store sc1), the wave waits until they are acknowledged, and
        // thread 0 counts the workgroup; the last workgroup reads the roots with
        // device-coherent loads.  The tree's own copy of the roots (plain stores) is read by
        // later launches only.
        if (t < 8) __hip_atomic_store(&a.xr[8 * blockIdx.x + t], in[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t < 64) {
            __builtin_amdgcn_s_waitcnt(0);
            if (t == 0) {
                const uint32_t b = blockIdx.x, g = b % TREE_CNT_GROUPS, ng = nb < TREE_CNT_GROUPS ? nb : TREE_CNT_GROUPS;
                const uint32_t gsize = (nb - g + TREE_CNT_GROUPS - 1) / TREE_CNT_GROUPS;
                uint32_t lst = 0;
                if (atomicAdd(&a.cnt[TREE_CNT_STRIDE * g], 1u) == gsize - 1)
                    lst = atomicAdd(&a.cnt[TREE_CNT_STRIDE * TREE_CNT_GROUPS], 1u) == ng - 1 ? 1u : 0u;
                last = lst;
            }
        }
        __syncthreads();
        if (!last) return;
        // Acquire: the loads of the other workgroups' roots below are ordered after the counter
        // update that observed their arrival (each publisher's roots were acknowledged, s_waitcnt,
        // before its count).  One fence per launch, in the last workgroup only.
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (t <= TREE_CNT_GROUPS) atomicExch(&a.cnt[TREE_CNT_STRIDE * t], 0u);
        if (a.ch && t < 4) {
#pragma unroll
            for (int k = 0; k < 8; k++) dg[k] = a.ch->digest[k];
        }
        // top levels from the nb subtree roots
        for (uint32_t i = t; i < 8 * nb; i += TL_THREADS)
            hA[i] = __hip_atomic_load(&a.xr[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        in = hA;
        out = hB;
        for (uint32_t c = nb >> 1; c >= 1; c >>= 1) {
            lev--;
            tr_level<TL_THREADS>(in, out, c, a.tree + (((uint64_t)1 << lev) - 1) * 32, 0, a.node_seed, t, rlo, rhi);
            uint32_t* x = in;
            in = out;
            out = x;
        }
    } else if (a.ch && t < 4) {
#pragma unroll
        for (int k = 0; k < 8; k++) dg[k] = a.ch->digest[k];
    }
    // quad 0 holds the root
    if (t < 4) {
        if (a.ch) {
            (void)tr_chan_step(L, q, dg, rlo, rhi, a.ch, a.root_out);
        } else if (a.root_out) {
            a.root_out[q] = rlo;
            a.root_out[4 + q] = rhi;
        }
    }
}

// fri_tail: the inner layers below the committed line top_log (<= FRI_TAIL_MAX_LOG + 1) run
// in one workgroup with the current line and hash level resident in LDS and the channel
// digest in registers of quad 0: per layer the fold into the next line, its leaf hashes,
// the parent levels (layers.zig:99-140) and the channel step on the root, separated by
// workgroup barriers instead of kernel boundaries (prover/fri.zig:377-621 runs the same
// sequence per layer).  The last fold (into last_log) stores the line only.
__global__ __launch_bounds__(TAIL_THREADS) void fri_tail_kernel(const FriTailArgs a) {
    // The per-log pointer arrays are read from the kernel arguments with the (uniform) log as
    // index and proof blockIdx.x's offset added at each use: a local copy of the arguments
    // indexed by a runtime log would live in scratch memory, and every scratch load's wait
    // would also wait for the tree stores of the previous level.
    const uint64_t boff = (uint64_t)blockIdx.x * a.bstride;
    const auto at = [boff](auto* p) { return (decltype(p))((uint8_t*)p + boff); };
    DevChannel* __restrict__ ch = at(a.ch);
    const B2State leaf_seed = a.leaf_seed, node_seed = a.node_seed;
    __shared__ __align__(16) uint32_t hA[TAIL_THREADS * 8];
    __shared__ __align__(16) uint32_t hB[TAIL_THREADS / 2 * 8];
    __shared__ uint4 vs[TAIL_THREADS];  // current line
    __shared__ uint32_t sa[4];          // alpha
    const uint32_t t = threadIdx.x, q = t & 3;
    const QuadLane L(q);
    uint32_t l = a.top_log;
    qm31 alpha = qload(ch->alpha);
    uint32_t dg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (t < 4) {
#pragma unroll
        for (int k = 0; k < 8; k++) dg[k] = ch->digest[k];
    }
    bool first = true;
    for (;;) {
        const uint32_t n = 1u << (l - 1);
        const uint32_t* inv_x = a.itw_half - ((size_t)1 << l);
        qm31 r = qzero();
        if (t < n) {
            const uint4* __restrict__ src = at(a.lines[l]);
            const uint4 x = first ? src[2 * t] : vs[2 * t];
            const uint4 y = first ? src[2 * t + 1] : vs[2 * t + 1];
            r = tr_fold(x, y, alpha, inv_x[t]);
        }
        first = false;
        __syncthreads();
        l--;
        const uint4 rv = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
        if (t < n) {
            vs[t] = rv;
            at(a.lines[l])[t] = rv;
        }
        if (l == a.last_log) break;
        uint8_t* tree = at(a.trees[l]);
        uint8_t* leaves = tree + (((size_t)1 << l) - 1) * 32;
        if (n <= TR_QUAD_MAX) {
            __syncthreads();
            for (uint32_t p = t >> 2; p < n; p += TAIL_THREADS / 4) {
                const uint4 v = vs[p];
                uint32_t hlo, hhi;
                tr_leaf_quad(L, leaf_seed, qm31{{v.x, v.y}, {v.z, v.w}}, hlo, hhi);
                tr_put_quad(hA + 8 * p, leaves + (size_t)p * 32, q, hlo, hhi);
            }
        } else if (t < n) {
            uint32_t h[8];
            tr_leaf_lane(leaf_seed, r, h);
            tr_put_lane(hA + 8 * t, leaves + (size_t)t * 32, h);
        }
        __syncthreads();
        uint32_t rlo = 0, rhi = 0;
        uint32_t *in = hA, *out = hB;
        uint32_t lev = l;
        for (uint32_t c = n >> 1; c >= 1; c >>= 1) {
            lev--;
            tr_level<TAIL_THREADS>(in, out, c, tree + (((size_t)1 << lev) - 1) * 32, 0, node_seed, t, rlo, rhi);
            uint32_t* x = in;
            in = out;
            out = x;
        }
        if (t < 4) {
            const qm31 na = tr_chan_step(L, q, dg, rlo, rhi, ch, a.roots[l] ? at(a.roots[l]) : nullptr);
            if (t == 0) qstore(na, sa);
        }
        __syncthreads();
        alpha = qload(sa);
    }
}

// Hashes per workgroup (log) for a tree reduction from level k0: the whole tree in one
// workgroup up to 2^8 leaves; above, the levels split about evenly between the workgroups'
// subtrees and the last workgroup's top, with at most 2^8 workgroups (one per CU: a second
// workgroup on a CU halves the issue rate of both) and at most 2^10 hashes per workgroup.
static uint32_t tree_chunk_log(uint32_t k0) { return k0 <= 8 ? k0 : std::min(10u, std::max(k0 / 2, k0 - 8)); }

template <int SRC>
static void launch_tree_layer(Ctx& c, FriLayerArgs& a, const std::string& cls, double bytes, uint32_t batch,
                              uint64_t bstride, uint32_t* d_cnt) {
    SR_CHECK(a.k0 >= 1 && a.k0 <= TREE_LAYER_MAX_LOG, "tree layer: level out of range");
    SR_CHECK(batch >= 1 && batch <= 65535, "tree layer: batch size");
    // one tree: spread over the CUs; a batch fills the device with its trees, so each workgroup
    // takes as many hashes as it can hold (fewer workgroups, shorter top reductions)
    a.s = batch > 1 ? std::min(a.k0, 10u) : tree_chunk_log(a.k0);
    const uint32_t nb = 1u << (a.k0 - a.s);
    SR_CHECK(nb <= TREE_MAX_BLOCKS && a.s >= 1 && a.s <= 10, "tree layer: chunking");
    // one tree: the context's counters; a batch: the caller's, one region per tree
    SR_CHECK(batch == 1 || d_cnt, "tree layer: a batch needs its own completion counters");
    uint32_t* cnt = batch > 1 ? d_cnt : c.d_tree_cnt;
    SR_CHECK(cnt, "tree layer: context without completion counters");
    a.cnt = cnt;
    a.xr = cnt + TREE_XR_OFFSET;
    a.leaf_seed = c.leaf_seed;
    a.node_seed = c.node_seed;
    a.bstride = bstride;
    KTimer kt(c, cls, batch * bytes);
    hipLaunchKernelGGL(tree_layer_kernel<SRC>, dim3(nb, batch), dim3(TL_THREADS), 0, c.stream, a);
    SR_HIP(hipGetLastError());
}

void launch_tree_top(Ctx& c, DeviceTree& t, uint32_t k0, DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch,
                     uint64_t bstride, uint32_t* d_cnt) {
    SR_CHECK(k0 <= t.max_log, "tree top: level above the leaves");
    FriLayerArgs a{};
    a.tree = t.layers;
    a.k0 = k0;
    a.ch = d_ch;
    a.root_out = (uint32_t*)d_root_out;
    launch_tree_layer<FRI_SRC_HASHES>(c, a, "merkle_top L=" + std::to_string(k0) + (d_ch ? " +channel" : ""),
                                      32.0 * (double)((2ull << k0) - 1), batch, bstride, d_cnt);
}

void launch_fri_commit_line(Ctx& c, const uint32_t* d_src, uint32_t log_src, uint32_t* d_dst, DeviceTree& t,
                            DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch, uint64_t bstride, uint32_t* d_cnt) {
    const uint32_t l = log_src - 1;
    SR_CHECK(log_src >= 1 && t.max_log == l && d_ch, "fri commit: layer shape");
    if (l == 0 || l > TREE_LAYER_MAX_LOG) {
        SR_CHECK(batch == 1, "fri commit: batches of layers above 2^16");
        launch_fold_line(c, d_src, log_src, qzero(), d_dst, t.layer(l), d_ch);
        launch_merkle_parents(c, t, d_ch, d_root_out);
        return;
    }
    c.ensure_tower(log_src + 1);
    FriLayerArgs a{};
    a.src_line = (const uint4*)d_src;
    a.inv = c.d_itw + ((1u << (c.tower_log - 1)) - (1u << log_src));
    a.dst_line = (uint4*)d_dst;
    a.tree = t.layers;
    a.k0 = l;
    a.ch = d_ch;
    a.root_out = (uint32_t*)d_root_out;
    const double n = (double)(1ull << l);
    launch_tree_layer<FRI_SRC_LINE>(c, a, "fri_layer", (32.0 + 4.0 + 16.0) * n + 32.0 * (2.0 * n - 1.0), batch,
                                    bstride, d_cnt);
}

void launch_fri_commit_circle(Ctx& c, const uint32_t* const* coords, uint32_t log_src, uint32_t* d_dst,
                              DeviceTree& t, DevChannel* d_ch, uint8_t* d_root_out, uint32_t batch, uint64_t bstride,
                              uint32_t* d_cnt) {
    const uint32_t l = log_src - 1;
    SR_CHECK(log_src >= 1 && t.max_log == l && d_ch, "fri commit: layer shape");
    if (l == 0 || l > TREE_LAYER_MAX_LOG) {
        SR_CHECK(batch == 1, "fri commit: batches of layers above 2^16");
        launch_fold_circle_into_line(c, coords, log_src, qzero(), d_dst, false, t.layer(l), d_ch);
        launch_merkle_parents(c, t, d_ch, d_root_out);
        return;
    }
    FriLayerArgs a{};
    for (int k = 0; k < 4; k++) a.coord[k] = coords[k];
    a.inv = c.inv_circle_y(log_src);
    a.dst_line = (uint4*)d_dst;
    a.tree = t.layers;
    a.k0 = l;
    a.ch = d_ch;
    a.root_out = (uint32_t*)d_root_out;
    const double n = (double)(1ull << l);
    launch_tree_layer<FRI_SRC_CIRCLE>(c, a, "fri_layer_circle", (32.0 + 4.0 + 16.0) * n + 32.0 * (2.0 * n - 1.0),
                                      batch, bstride, d_cnt);
}

void launch_fri_tail(Ctx& c, FriTailArgs& a, uint32_t batch) {
    SR_CHECK(a.top_log <= FRI_TAIL_MAX_LOG + 1 && a.top_log > a.last_log, "fri_tail: layer range");
    SR_CHECK(batch >= 1 && batch <= 65535, "fri_tail: batch size");
    c.ensure_tower(a.top_log + 1);
    // inverse twiddles of halfOdds(l) live at [2^(M-1) - 2^l, 2^(M-1) - 2^(l-1)) of the inverse tower
    a.itw_half = c.d_itw + (1u << (c.tower_log - 1));
    a.leaf_seed = c.leaf_seed;
    a.node_seed = c.node_seed;
    KTimer kt(c, "fri_tail", 0.0);
    hipLaunchKernelGGL(fri_tail_kernel, dim3(batch), dim3(TAIL_THREADS), 0, c.stream, a);
    SR_HIP(hipGetLastError());
}

void launch_fold_line(Ctx& c, const uint32_t* d_src, uint32_t log_src, qm31 alpha, uint32_t* d_dst,
                      uint8_t* d_leaf_out, const DevChannel* d_alpha) {
    SR_CHECK(log_src >= 1, "fold_line: source too short");
    c.ensure_tower(log_src + 1);
    // inverse twiddle layer of halfOdds(log_src): length 2^(log_src-1) at 2^(M-1) - 2^log_src.
    const uint32_t* inv_x = c.d_itw + ((1u << (c.tower_log - 1)) - (1u << log_src));
    uint32_t n = 1u << (log_src - 1);
    KTimer kt(c, "fri_fold_line", (32.0 + 4.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n);
    hipLaunchKernelGGL(fold_line_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, (const uint4*)d_src, log_src,
                       inv_x, alpha, d_alpha, (uint4*)d_dst, d_leaf_out, c.leaf_seed);
    SR_HIP(hipGetLastError());
}

void launch_fold_circle_into_line(Ctx& c, const uint32_t* const* coords, uint32_t log_src, qm31 alpha,
                                  uint32_t* d_dst, bool accumulate, uint8_t* d_leaf_out, const DevChannel* d_alpha,
                                  uint32_t batch, uint64_t bstride) {
    SR_CHECK(log_src >= 1, "fold_circle: source too short");
    SR_CHECK(batch >= 1 && batch <= 65535, "fold_circle: batch size");
    const uint32_t* inv_y = c.inv_circle_y(log_src);
    uint32_t n = 1u << (log_src - 1);
    KTimer kt(c, "fri_fold_circle", batch * (32.0 + 4.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n);
    hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, coords[0], coords[1],
                       coords[2], coords[3], log_src, inv_y, alpha, qsqr(alpha), d_alpha, accumulate,
                       (uint4*)d_dst, d_leaf_out, c.leaf_seed, bstride);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
