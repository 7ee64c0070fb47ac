// Small kernels: proof-of-work grind, decommitment gather, batch inverse,
// bit reversal, accumulation and QM31 AoS -> coordinate columns.
#include "runtime.hpp"

namespace sr {

// PoW (src/core/channel/blake2s.zig:135-166,289-300): H(prefix32 || LE64(nonce)),
// one 40-byte final block; valid when the first 16 bytes, read as a
// little-endian 128-bit integer, have >= bits trailing zeros.  The lowest valid
// nonce of the launch range is kept with a 64-bit atomic min (vector atomic).
__device__ __forceinline__ void pow_try(const B2State& init, uint4 p0, uint4 p1, uint32_t bits, uint64_t nonce,
                                        unsigned long long* __restrict__ best) {
    uint32_t m[16] = {p0.x, p0.y, p0.z, p0.w, p1.x, p1.y, p1.z, p1.w,
                      (uint32_t)nonce, (uint32_t)(nonce >> 32), 0, 0, 0, 0, 0, 0};
    uint32_t h[8];
#pragma unroll
    for (int k = 0; k < 8; k++) h[k] = init.h[k];
    b2_compress(h, m, 40, true);
    uint32_t tz;
    if (h[0]) tz = __builtin_ctz(h[0]);
    else if (h[1]) tz = 32 + __builtin_ctz(h[1]);
    else if (h[2]) tz = 64 + __builtin_ctz(h[2]);
    else if (h[3]) tz = 96 + __builtin_ctz(h[3]);
    else tz = 128;
    if (tz >= bits) atomicMin(best, (unsigned long long)nonce);
}
__global__ __launch_bounds__(256) void pow_kernel(B2State init, uint4 p0, uint4 p1, uint32_t bits, uint64_t base,
                                                  uint64_t count, unsigned long long* __restrict__ best) {
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    pow_try(init, p0, p1, bits, base + i, best);
}
// Prefix blockIdx.y of a batch: its 8 words at prefix + 8 y, its best nonce at best[y].  A
// prefix whose nonce an earlier range found (best < base) is skipped; a value this launch
// wrote (>= base) does not skip, so the lowest valid nonce is found either way.
__global__ __launch_bounds__(256) void pow_batch_kernel(B2State init, const uint4* __restrict__ prefix, uint32_t bits,
                                                        uint64_t base, uint64_t count,
                                                        unsigned long long* __restrict__ best) {
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    if (best[blockIdx.y] < base) return;
    const uint4 p0 = prefix[2 * blockIdx.y], p1 = prefix[2 * blockIdx.y + 1];
    pow_try(init, p0, p1, bits, base + i, best + blockIdx.y);
}
void launch_pow_grind_batch(Ctx& c, const uint32_t* d_prefix, uint32_t batch, uint32_t bits, uint64_t base,
                            uint64_t count, unsigned long long* d_best) {
    SR_CHECK(batch >= 1 && batch <= 65535, "pow: batch size");
    B2State st;
    b2_init(st.h);
    const uint64_t blocks = (count + 255) / 256;
    SR_CHECK(blocks < (1ull << 31), "pow: launch range too large");
    KTimer kt(c, "pow_grind", 0.0);
    hipLaunchKernelGGL(pow_batch_kernel, dim3((uint32_t)blocks, batch), dim3(256), 0, c.stream, st,
                       (const uint4*)d_prefix, bits, base, count, d_best);
    SR_HIP(hipGetLastError());
}

uint64_t launch_pow_grind(Ctx& c, const uint8_t prefix[32], uint32_t bits, uint64_t base, uint64_t count) {
    DBuf best(&c, 8);
    uint64_t init = UINT64_MAX;
    SR_HIP(hipMemcpyAsync(best.p, c.stage(&init, 8), 8, hipMemcpyHostToDevice, c.stream));
    B2State st;
    b2_init(st.h);
    uint32_t w[8];
    memcpy(w, prefix, 32);
    uint64_t blocks = (count + 255) / 256;
    SR_CHECK(blocks < (1ull << 31), "pow: launch range too large");
    KTimer kt(c, "pow_grind", 0.0);
    hipLaunchKernelGGL(pow_kernel, dim3((uint32_t)blocks), dim3(256), 0, c.stream, st,
                       make_uint4(w[0], w[1], w[2], w[3]), make_uint4(w[4], w[5], w[6], w[7]), bits, base, count,
                       best.as<unsigned long long>());
    SR_HIP(hipGetLastError());
    uint64_t* h = (uint64_t*)c.pinned(8);
    SR_HIP(hipMemcpyAsync(h, best.p, 8, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    return *h;
}

// Decommitment gather: out[i] = word at absolute device address addr[i].
__global__ void gather_kernel(const uint64_t* __restrict__ addr, uint64_t n, uint32_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *(gcu32*)(uintptr_t)addr[i];
}
void launch_gather(Ctx& c, const uint64_t* d_addr, uint64_t n, uint32_t* d_out) {
    if (!n) return;
    hipLaunchKernelGGL(gather_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream, d_addr, n, d_out);
    SR_HIP(hipGetLastError());
}

// Field inverse per element (fields/mod.zig:16 batchInverse; the inverse is unique).
__global__ void inverse_kernel(const uint32_t* __restrict__ in, uint64_t n, uint32_t* __restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = minv(in[i]);
}
void launch_batch_inverse(Ctx& c, const uint32_t* d_in, uint64_t n, uint32_t* d_out) {
    if (!n) return;
    hipLaunchKernelGGL(inverse_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream, d_in, n, d_out);
    SR_HIP(hipGetLastError());
}

// In-place bit reversal of 2^log elements of `wpe` words (core/utils.zig:20-34).
__global__ void bitrev_kernel(uint32_t* __restrict__ v, uint32_t log_size, uint32_t wpe) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << log_size)) return;
    uint32_t j = bitrev32(i, log_size);
    if (j <= i) return;
    for (uint32_t k = 0; k < wpe; k++) {
        uint32_t t = v[(uint64_t)i * wpe + k];
        v[(uint64_t)i * wpe + k] = v[(uint64_t)j * wpe + k];
        v[(uint64_t)j * wpe + k] = t;
    }
}
void launch_bit_reverse(Ctx& c, uint32_t* d_vals, uint32_t log_size, uint32_t wpe) {
    uint32_t n = 1u << log_size;
    hipLaunchKernelGGL(bitrev_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, d_vals, log_size, wpe);
    SR_HIP(hipGetLastError());
}

// AccumulationOps.accumulate (backend/accumulation_ops.zig): column-wise add.
__global__ void accumulate_kernel(uint32_t* __restrict__ acc, const uint32_t* __restrict__ src, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = madd(acc[i], src[i]);
}
void launch_accumulate(Ctx& c, uint32_t* d_acc, const uint32_t* d_src, uint64_t n_words) {
    if (!n_words) return;
    hipLaunchKernelGGL(accumulate_kernel, dim3((uint32_t)((n_words + 255) / 256)), dim3(256), 0, c.stream, d_acc,
                       d_src, n_words);
    SR_HIP(hipGetLastError());
}

__global__ void coords_kernel(const uint4* __restrict__ q, uint64_t n, uint32_t* __restrict__ c0,
                              uint32_t* __restrict__ c1, uint32_t* __restrict__ c2, uint32_t* __restrict__ c3) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint4 v = q[i];
    c0[i] = v.x;
    c1[i] = v.y;
    c2[i] = v.z;
    c3[i] = v.w;
}
void launch_qm31_to_coords(Ctx& c, const uint32_t* d_qm31, uint64_t n, uint32_t* const* coords) {
    if (!n) return;
    hipLaunchKernelGGL(coords_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream,
                       (const uint4*)d_qm31, n, coords[0], coords[1], coords[2], coords[3]);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
