// FRI folds on gfx950 (src/core/fri/folding.zig:316-380,558-602).
//
// fold_line:   dst[i] = (f0 + f1) + alpha * (f0 - f1) * x_i^-1, (f0, f1) = src[2i], src[2i+1],
//              x_i = LineDomain(halfOdds(log)).at(bitrev(i)).x.  x_i^-1 is read from the
//              inverse twiddle tower: its layer for halfOdds(log) holds exactly these values.
// fold_circle: dst[i] = dst[i] * alpha^2 + (f0 + f1) + alpha * (f0 - f1) * y_i^-1 with
//              y_i = halfOdds(log-1).at(bitrev(i)).y (resident table per log size).
// Both optionally hash the folded value as the next layer's Merkle leaf
// (prefix || 4 LE words, one compression), fusing fold and leaf hashing.
#include "runtime.hpp"

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
                                                        uint4* __restrict__ dst, uint8_t* __restrict__ leaves,
                                                        B2State seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << (log_src - 1))) return;
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
                                                          bool accumulate, uint4* __restrict__ dst,
                                                          uint8_t* __restrict__ leaves, B2State seed) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << (log_src - 1))) return;
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

void launch_fold_line(Ctx& c, const uint32_t* d_src, uint32_t log_src, qm31 alpha, uint32_t* d_dst,
                      uint8_t* d_leaf_out) {
    SR_CHECK(log_src >= 1, "fold_line: source too short");
    c.ensure_tower(log_src + 1);
    // inverse twiddle layer of halfOdds(log_src): length 2^(log_src-1) at 2^(M-1) - 2^log_src.
    const uint32_t* inv_x = c.d_itw + ((1u << (c.tower_log - 1)) - (1u << log_src));
    uint32_t n = 1u << (log_src - 1);
    KTimer kt(c, "fri_fold_line", (32.0 + 4.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n);
    hipLaunchKernelGGL(fold_line_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, (const uint4*)d_src, log_src,
                       inv_x, alpha, (uint4*)d_dst, d_leaf_out, c.leaf_seed);
    SR_HIP(hipGetLastError());
}

void launch_fold_circle_into_line(Ctx& c, const uint32_t* const* coords, uint32_t log_src, qm31 alpha,
                                  uint32_t* d_dst, bool accumulate, uint8_t* d_leaf_out) {
    SR_CHECK(log_src >= 1, "fold_circle: source too short");
    const uint32_t* inv_y = c.inv_circle_y(log_src);
    uint32_t n = 1u << (log_src - 1);
    KTimer kt(c, "fri_fold_circle", (32.0 + 4.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n);
    hipLaunchKernelGGL(fold_circle_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, coords[0], coords[1],
                       coords[2], coords[3], log_src, inv_y, alpha, qsqr(alpha), accumulate, (uint4*)d_dst,
                       d_leaf_out, c.leaf_seed);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
