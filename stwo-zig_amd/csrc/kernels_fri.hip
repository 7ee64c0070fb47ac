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
// fri_tail: the inner layers of log <= FRI_TAIL_MAX_LOG run in one workgroup,
// with the current line and hash level resident in LDS: per layer, the parent
// levels (layers.zig:99-140), the channel step on the root, the fold into the
// next line and its leaf hashes, separated by workgroup barriers instead of
// kernel boundaries (prover/fri.zig:377-621 runs the same sequence per layer).
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
    qm31 r = qadd(qadd(f0, f1), qmul(alpha, qmulm(qsub(f0, f1), inv_x[i])));
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
                                                          B2State seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << (log_src - 1))) return;
    if (dch) {
        alpha = qload(dch->alpha);
        alpha2 = qload(dch->alpha2);
    }
    uint2 w0 = ((const uint2*)c0)[i], w1 = ((const uint2*)c1)[i], w2 = ((const uint2*)c2)[i],
          w3 = ((const uint2*)c3)[i];
    qm31 f0 = {{w0.x, w1.x}, {w2.x, w3.x}}, f1 = {{w0.y, w1.y}, {w2.y, w3.y}};
    qm31 r = qadd(qadd(f0, f1), qmul(alpha, qmulm(qsub(f0, f1), inv_y[i])));
    if (accumulate) {
        uint4 d = dst[i];
        r = qadd(qmul(qm31{{d.x, d.y}, {d.z, d.w}}, alpha2), r);
    }
    dst[i] = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
    if (leaves) leaf_qm31(leaves, i, seed, r);
}

__device__ __forceinline__ void tail_node_hash(const B2State& seed, const uint32_t* l, const uint32_t* r,
                                               uint32_t h[8]) {
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

__device__ __forceinline__ void tail_store_hash(uint8_t* __restrict__ out, uint32_t i, const uint32_t h[8]) {
    uint4* o = (uint4*)(out + (size_t)i * 32);
    o[0] = make_uint4(h[0], h[1], h[2], h[3]);
    o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

constexpr int FRI_TAIL_THREADS = 1 << FRI_TAIL_MAX_LOG;

__global__ __launch_bounds__(FRI_TAIL_THREADS) void fri_tail_kernel(FriTailArgs a) {
    __shared__ uint32_t hs[FRI_TAIL_THREADS * 8];  // current hash level
    __shared__ uint4 vs[FRI_TAIL_THREADS];         // current line
    __shared__ uint32_t sa[4];                     // alpha
    const uint32_t t = threadIdx.x;
    uint32_t l = a.top_log;
    if (t < (1u << l)) {
        const uint4* lf = (const uint4*)(a.trees[l] + (((size_t)1 << l) - 1) * 32) + 2 * t;
        uint4 x = lf[0], y = lf[1];
        hs[t * 8 + 0] = x.x, hs[t * 8 + 1] = x.y, hs[t * 8 + 2] = x.z, hs[t * 8 + 3] = x.w;
        hs[t * 8 + 4] = y.x, hs[t * 8 + 5] = y.y, hs[t * 8 + 6] = y.z, hs[t * 8 + 7] = y.w;
        vs[t] = a.lines[l][t];
    }
    __syncthreads();
    while (l > a.last_log) {
        // parent levels of the tree over this line, down to its root; levels of
        // at most FRI_TAIL_THREADS/8 nodes use one quad per compression
        uint32_t rlo = 0, rhi = 0;
        for (int32_t lev = (int32_t)l - 1; lev >= 0; lev--) {
            const uint32_t cnt = 1u << lev;
            uint8_t* out = a.trees[l] + (((size_t)1 << lev) - 1) * 32;
            if (cnt > FRI_TAIL_THREADS / 8) {
                uint32_t h[8];
                const bool act = t < cnt;
                if (act) tail_node_hash(a.node_seed, &hs[16 * t], &hs[16 * t + 8], h);
                __syncthreads();
                if (act) {
#pragma unroll
                    for (int i = 0; i < 8; i++) hs[8 * t + i] = h[i];
                    tail_store_hash(out, t, h);
                }
            } else {
                const uint32_t q = t & 3, j = t >> 2;
                const QuadLane L(q);
                const bool act = j < cnt;
                if (act) {
                    uint32_t m[16];
#pragma unroll
                    for (int i = 0; i < 16; i++) m[i] = hs[16 * j + i];
                    rlo = L.sel(a.node_seed.h[0], a.node_seed.h[1], a.node_seed.h[2], a.node_seed.h[3]);
                    rhi = L.sel(a.node_seed.h[4], a.node_seed.h[5], a.node_seed.h[6], a.node_seed.h[7]);
                    b2_compress_quad(L, rlo, rhi, m, 128, true);
                }
                __syncthreads();
                if (act) {
                    hs[8 * j + q] = rlo;
                    hs[8 * j + 4 + q] = rhi;
                    uint32_t* o = (uint32_t*)(out + (size_t)j * 32);
                    o[q] = rlo;
                    o[4 + q] = rhi;
                }
            }
            __syncthreads();
        }
        // lanes 0..3 hold the root (last level ran in quad mode)
        if (t < 4) {
            chan_mix_root_draw_quad(a.ch, rlo, rhi, a.roots[l]);
            if (t == 0) {
#pragma unroll
                for (int i = 0; i < 4; i++) sa[i] = a.ch->alpha[i];
            }
        }
        __syncthreads();
        // fold line l into line l-1 (fold_line_kernel above), hashing the leaves of the next tree
        const qm31 alpha = qload(sa);
        const uint32_t n = 1u << (l - 1);
        qm31 r;
        if (t < n) {
            uint4 x = vs[2 * t], y = vs[2 * t + 1];
            qm31 f0 = {{x.x, x.y}, {x.z, x.w}}, f1 = {{y.x, y.y}, {y.z, y.w}};
            const uint32_t* inv_x = a.itw_half - ((size_t)1 << l);
            r = qadd(qadd(f0, f1), qmul(alpha, qmulm(qsub(f0, f1), inv_x[t])));
        }
        __syncthreads();
        l--;
        if (t < n) {
            uint4 rv = make_uint4(r.c0.a, r.c0.b, r.c1.a, r.c1.b);
            vs[t] = rv;
            a.lines[l][t] = rv;
            if (l > a.last_log) {
                uint32_t m[16] = {r.c0.a, r.c0.b, r.c1.a, r.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
                uint32_t h[8];
#pragma unroll
                for (int k = 0; k < 8; k++) h[k] = a.leaf_seed.h[k];
                b2_compress(h, m, 80, true);
#pragma unroll
                for (int i = 0; i < 8; i++) hs[8 * t + i] = h[i];
                tail_store_hash(a.trees[l] + (((size_t)1 << l) - 1) * 32, t, h);
            }
        }
        __syncthreads();
    }
}

void launch_fri_tail(Ctx& c, FriTailArgs& a) {
    SR_CHECK(a.top_log <= FRI_TAIL_MAX_LOG && a.top_log > a.last_log, "fri_tail: layer range");
    c.ensure_tower(a.top_log + 1);
    // inverse twiddles of halfOdds(l) live at [2^(M-1) - 2^l, 2^(M-1) - 2^(l-1)) of the inverse tower
    a.itw_half = c.d_itw + (1u << (c.tower_log - 1));
    a.leaf_seed = c.leaf_seed;
    a.node_seed = c.node_seed;
    KTimer kt(c, "fri_tail", 0.0);
    hipLaunchKernelGGL(fri_tail_kernel, dim3(1), dim3(FRI_TAIL_THREADS), 0, c.stream, a);
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
                                  uint32_t* d_dst, bool accumulate, uint8_t* d_leaf_out, const DevChannel* d_alpha) {
    SR_CHECK(log_src >= 1, "fold_circle: source too short");
    const uint32_t* inv_y = c.inv_circle_y(log_src);
    uint32_t n = 1u << (log_src - 1);
    KTimer kt(c, "fri_fold_circle", (32.0 + 4.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n);
    hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, coords[0], coords[1],
                       coords[2], coords[3], log_src, inv_y, alpha, qsqr(alpha), d_alpha, accumulate,
                       (uint4*)d_dst, d_leaf_out, c.leaf_seed);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
