// LogUp / GKR column operations on gfx950.
//
// Fractions (src/core/fraction.zig:9-73), the cumulative column of a LogUp
// interaction (src/frontends/riscv/air/logup.zig:74-91) and the GKR backend
// hooks genEqEvals / nextLayer / sumAsPolyInFirstVariable
// (src/backend/gkr_ops.zig:1-23; src/prover/lookups/gkr_circuit.zig:231-674,
// mle.zig:80-95), plus the XOR-lookup AIR of the LogUp workload: a, b uniform
// 8-bit from splitmix64, c = a ^ b, a 2^16-entry table (x, y, x ^ y) with
// multiplicities, denominators z - (a + b*alpha + c*alpha^2).
//
// All arithmetic is exact in M31/QM31, so every reduction order gives the same
// words: sums are taken per coordinate in 64-bit integers (atomic adds across
// blocks), the prefix scan adds QM31 values in any tree order, and inverses are
// per element (the inverse is unique).
#include <mutex>

#include "runtime.hpp"

namespace sr {

constexpr int LG_THREADS = 256;

__device__ __forceinline__ uint64_t lg_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}
__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ __launch_bounds__(LG_THREADS) void xor_trace_kernel(uint32_t log_size, uint64_t seed, uint32_t* __restrict__ a,
                                                               uint32_t* __restrict__ b, uint32_t* __restrict__ c,
                                                               uint32_t* __restrict__ mult) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1ull << log_size)) return;
    uint64_t w = lg_splitmix64(seed ^ (i * 0xD1B54A32D192ED03ull));
    uint32_t x = (uint32_t)(w & 0xff), y = (uint32_t)((w >> 8) & 0xff);
    a[i] = x;
    b[i] = y;
    c[i] = x ^ y;
    atomicAdd(&mult[(x << 8) | y], 1u);
}

// den[i] = z - (a + b alpha + c alpha^2); table rows (x = (i >> 8) & 255, y = i & 255, x ^ y) when
// a == nullptr (rows beyond 2^16 repeat the table and carry zero multiplicity).  With zs set,
// z = zs[0] and alpha = zs[1] are read from device memory (drawn by a device transcript step).
__global__ __launch_bounds__(LG_THREADS) void logup_denoms_kernel(const uint32_t* __restrict__ a,
                                                                  const uint32_t* __restrict__ b,
                                                                  const uint32_t* __restrict__ c, uint64_t n, qm31 z,
                                                                  qm31 alpha, qm31 alpha2, const uint4* __restrict__ zs,
                                                                  uint4* __restrict__ den) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (zs) {
        const uint4 zv = zs[0], av = zs[1];
        z = {{zv.x, zv.y}, {zv.z, zv.w}};
        alpha = {{av.x, av.y}, {av.z, av.w}};
        alpha2 = qsqr(alpha);
    }
    uint32_t x, y, r;
    if (a) {
        x = a[i];
        y = b[i];
        r = c[i];
    } else {
        x = (uint32_t)((i >> 8) & 0xff);
        y = (uint32_t)(i & 0xff);
        r = x ^ y;
    }
    qm31 comb = qadd(qadd(qfromm(x), qmulm(alpha, y)), qmulm(alpha2, r));
    qm31 d = qsub(z, comb);
    den[i] = make_uint4(d.c0.a, d.c0.b, d.c1.a, d.c1.b);
}

__device__ __forceinline__ qm31 ld4(const uint4* p, uint64_t i) {
    uint4 v = p[i];
    return {{v.x, v.y}, {v.z, v.w}};
}
__device__ __forceinline__ void st4(uint4* p, uint64_t i, qm31 v) { p[i] = make_uint4(v.c0.a, v.c0.b, v.c1.a, v.c1.b); }

// ---- cumulative column: values num/den, inclusive scan in trace order ----
constexpr int SCAN_PER_THREAD = 8;
constexpr int SCAN_TILE = LG_THREADS * SCAN_PER_THREAD;  // elements per block

// Block-local inclusive scan of QM31 values held one per thread (Hillis-Steele in LDS).
__device__ __forceinline__ qm31 block_scan_incl(qm31 v, qm31* sh) {
    const uint32_t t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (uint32_t off = 1; off < LG_THREADS; off <<= 1) {
        qm31 add = t >= off ? sh[t - off] : qzero();
        __syncthreads();
        v = qadd(v, add);
        sh[t] = v;
        __syncthreads();
    }
    return v;
}

// num_kind: 0 = QM31 AoS, 1 = M31, 2 = ones.  Each block scans SCAN_TILE elements
// (8 consecutive per thread) and writes its total to block_sums.
__global__ __launch_bounds__(LG_THREADS) void logup_scan_tile_kernel(int num_kind, const uint32_t* __restrict__ num,
                                                                     const uint4* __restrict__ den, uint64_t n,
                                                                     uint4* __restrict__ sums,
                                                                     uint4* __restrict__ block_sums,
                                                                     uint32_t* __restrict__ zero_flag) {
    __shared__ qm31 sh[LG_THREADS];
    const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_PER_THREAD;
    qm31 loc[SCAN_PER_THREAD];
    qm31 run = qzero();
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        const uint64_t i = base + k;
        qm31 v = qzero();
        if (i < n) {
            qm31 d = ld4(den, i);
            if (qiszero(d)) {
                atomicOr(zero_flag, 1u);
            } else {
                qm31 nu = num_kind == 0 ? ld4((const uint4*)num, i) : num_kind == 1 ? qfromm(num[i]) : qone();
                v = qmul(nu, qinv(d));
            }
        }
        run = qadd(run, v);
        loc[k] = run;
    }
    qm31 incl = block_scan_incl(run, sh);
    qm31 excl = qsub(incl, run);
#pragma unroll
    for (int k = 0; k < SCAN_PER_THREAD; k++) {
        const uint64_t i = base + k;
        if (i < n) st4(sums, i, qadd(excl, loc[k]));
    }
    if (threadIdx.x == LG_THREADS - 1) st4(block_sums, blockIdx.x, incl);
}

// Exclusive scan of the block totals by one block (looping over chunks), in place.
__global__ __launch_bounds__(LG_THREADS) void logup_scan_blocks_kernel(uint4* __restrict__ block_sums, uint32_t n_blocks,
                                                                       uint4* __restrict__ total) {
    __shared__ qm31 sh[LG_THREADS];
    qm31 carry = qzero();
    for (uint32_t start = 0; start < n_blocks; start += LG_THREADS) {
        const uint32_t i = start + threadIdx.x;
        qm31 v = i < n_blocks ? ld4(block_sums, i) : qzero();
        qm31 incl = block_scan_incl(v, sh);
        if (i < n_blocks) st4(block_sums, i, qadd(carry, qsub(incl, v)));
        __syncthreads();
        carry = qadd(carry, sh[LG_THREADS - 1]);
        __syncthreads();
    }
    if (threadIdx.x == 0) st4(total, 0, carry);
}

__global__ __launch_bounds__(LG_THREADS) void logup_scan_add_kernel(uint4* __restrict__ sums, uint64_t n,
                                                                    const uint4* __restrict__ block_prefix) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t blk = i / SCAN_TILE;
    if (blk == 0) return;
    st4(sums, i, qadd(ld4(sums, i), ld4(block_prefix, blk)));
}

// ---- GKR hooks ----
// out[i] = scale * prod_j (bit_{n-1-j}(i) ? y_j : 1 - y_j)  (gkr_circuit.zig:651-674).
__global__ __launch_bounds__(LG_THREADS) void gkr_eq_evals_kernel(const uint4* __restrict__ y, uint32_t n, qm31 scale,
                                                                  uint4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1ull << n)) return;
    qm31 acc = scale;
    for (uint32_t j = 0; j < n; j++) {
        qm31 yj = ld4(y, j);
        acc = qmul(acc, ((i >> (n - 1 - j)) & 1) ? yj : qsub(qone(), yj));
    }
    st4(out, i, acc);
}

__device__ __forceinline__ qm31 layer_num(int kind, const uint32_t* num, uint64_t i) {
    if (kind == GKR_LOGUP_GENERIC) return ld4((const uint4*)num, i);
    if (kind == GKR_LOGUP_MULTIPLICITIES) return qfromm(num[i]);
    return qone();
}

// nextLayer (gkr_circuit.zig:449-503): pairs (2i, 2i+1).
__global__ __launch_bounds__(LG_THREADS) void gkr_next_layer_kernel(int kind, const uint32_t* __restrict__ num,
                                                                    const uint4* __restrict__ den, uint64_t half,
                                                                    uint4* __restrict__ out_num,
                                                                    uint4* __restrict__ out_den) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    qm31 d0 = ld4(den, 2 * i), d1 = ld4(den, 2 * i + 1);
    if (kind == GKR_GRAND_PRODUCT) {
        st4(out_den, i, qmul(d0, d1));
        return;
    }
    qm31 n0 = layer_num(kind, num, 2 * i), n1 = layer_num(kind, num, 2 * i + 1);
    st4(out_num, i, qadd(qmul(d1, n0), qmul(d0, n1)));
    st4(out_den, i, qmul(d0, d1));
}

// sum_{i < n_terms} eq[i] * term_r(i) for r = 0 and r = 2 (gkr_circuit.zig:505-643), per
// coordinate in 64-bit: block partials, then one atomic add per word into acc[8].
__global__ __launch_bounds__(LG_THREADS) void gkr_sum_kernel(int kind, const uint4* __restrict__ eq,
                                                             const uint32_t* __restrict__ num,
                                                             const uint4* __restrict__ den, uint64_t nt, qm31 lambda,
                                                             unsigned long long* __restrict__ acc) {
    __shared__ unsigned long long red[8][LG_THREADS];
    uint64_t s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nt; i += (uint64_t)gridDim.x * blockDim.x) {
        qm31 t0, t2;
        if (kind == GKR_GRAND_PRODUCT) {
            qm31 a0 = ld4(den, 2 * i), a1 = ld4(den, 2 * i + 1);
            qm31 b0 = ld4(den, 2 * (nt + i)), b1 = ld4(den, 2 * (nt + i) + 1);
            t0 = qmul(a0, a1);
            t2 = qmul(qsub(qadd(b0, b0), a0), qsub(qadd(b1, b1), a1));
        } else {
            qm31 d00 = ld4(den, 2 * i), d01 = ld4(den, 2 * i + 1);
            qm31 d10 = ld4(den, 2 * (nt + i)), d11 = ld4(den, 2 * (nt + i) + 1);
            qm31 d20 = qsub(qadd(d10, d10), d00), d21 = qsub(qadd(d11, d11), d01);
            qm31 n00 = layer_num(kind, num, 2 * i), n01 = layer_num(kind, num, 2 * i + 1);
            qm31 n10 = layer_num(kind, num, 2 * (nt + i)), n11 = layer_num(kind, num, 2 * (nt + i) + 1);
            qm31 n20 = qsub(qadd(n10, n10), n00), n21 = qsub(qadd(n11, n11), n01);
            t0 = qadd(qadd(qmul(d01, n00), qmul(d00, n01)), qmul(qmul(d00, d01), lambda));
            t2 = qadd(qadd(qmul(d21, n20), qmul(d20, n21)), qmul(qmul(d20, d21), lambda));
        }
        qm31 e = ld4(eq, i);
        qm31 v0 = qmul(e, t0), v2 = qmul(e, t2);
        s[0] += v0.c0.a, s[1] += v0.c0.b, s[2] += v0.c1.a, s[3] += v0.c1.b;
        s[4] += v2.c0.a, s[5] += v2.c0.b, s[6] += v2.c1.a, s[7] += v2.c1.b;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (uint32_t st = LG_THREADS / 2; st > 0; st >>= 1) {
        if (threadIdx.x < st)
#pragma unroll
            for (int k = 0; k < 8; k++) red[k][threadIdx.x] += red[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < 8) atomicAdd(&acc[threadIdx.x], red[threadIdx.x][0]);
}

// Mle.fixFirstVariable: out[i] = r * (e[i + half] - e[i]) + e[i].
__global__ __launch_bounds__(LG_THREADS) void mle_fix_first_kernel(int base, const uint32_t* __restrict__ in,
                                                                   uint64_t half, qm31 r, uint4* __restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    qm31 e0 = base ? qfromm(in[i]) : ld4((const uint4*)in, i);
    qm31 e1 = base ? qfromm(in[i + half]) : ld4((const uint4*)in, i + half);
    st4(out, i, qadd(qmul(qsub(e1, e0), r), e0));
}

// Table columns of the LogUp workload on n >= 2^16 rows: mult[i] for i < 2^16 (0 beyond) and the
// GKR numerators -mult[i] mod p.
__global__ __launch_bounds__(LG_THREADS) void logup_table_cols_kernel(const uint32_t* __restrict__ mult16, uint64_t n,
                                                                      uint32_t* __restrict__ mpad,
                                                                      uint32_t* __restrict__ negm) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t m = i < (1u << 16) ? mult16[i] : 0u;
    mpad[i] = m;
    negm[i] = mneg(m);
}

static dim3 grid_for(uint64_t n) {
    uint64_t b = (n + LG_THREADS - 1) / LG_THREADS;
    SR_CHECK(b < (1ull << 31), "grid too large");
    return dim3((uint32_t)(b ? b : 1));
}

void launch_fill_u32(Ctx& c, uint32_t* d, uint64_t n, uint32_t v) {
    if (!n) return;
    hipLaunchKernelGGL(fill_u32_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, d, n, v);
    SR_HIP(hipGetLastError());
}

void launch_xor_trace(Ctx& c, uint32_t log_size, uint64_t seed, uint32_t* a, uint32_t* b, uint32_t* cc,
                      uint32_t* mult) {
    SR_CHECK(log_size <= 30, "xor trace: log size");
    launch_fill_u32(c, mult, 1u << 16, 0);
    uint64_t n = 1ull << log_size;
    KTimer kt(c, "xor_trace", 16.0 * n);
    hipLaunchKernelGGL(xor_trace_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, log_size, seed, a, b, cc, mult);
    SR_HIP(hipGetLastError());
}

void launch_logup_denominators(Ctx& c, const uint32_t* a, const uint32_t* b, const uint32_t* cc, uint64_t n, qm31 z,
                               qm31 alpha, uint32_t* den) {
    KTimer kt(c, "logup_denominators", (a ? 12.0 : 0.0) * n + 16.0 * n);
    hipLaunchKernelGGL(logup_denoms_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, a, b, cc, n, z, alpha,
                       qsqr(alpha), (const uint4*)nullptr, (uint4*)den);
    SR_HIP(hipGetLastError());
}

void launch_logup_denominators_dev(Ctx& c, const uint32_t* a, const uint32_t* b, const uint32_t* cc, uint64_t n,
                                   const uint32_t* d_z_alpha, uint32_t* den) {
    SR_CHECK(d_z_alpha, "denominators: challenges");
    KTimer kt(c, "logup_denominators", (a ? 12.0 : 0.0) * n + 16.0 * n);
    hipLaunchKernelGGL(logup_denoms_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, a, b, cc, n, qzero(), qzero(),
                       qzero(), (const uint4*)d_z_alpha, (uint4*)den);
    SR_HIP(hipGetLastError());
}

static uint64_t scan_blocks(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE; }

// The three scan launches; d_bs holds scan_blocks(n) block sums (16 B each).
static void launch_cumulative(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n,
                              uint32_t* sums, uint4* d_bs, uint4* d_total, uint32_t* d_flag) {
    SR_CHECK(n > 0, "cumulative column: empty");
    SR_CHECK(num_kind >= 0 && num_kind <= 2 && (num_kind == 2 || num), "cumulative column: numerators");
    const uint64_t n_blocks = scan_blocks(n);
    SR_CHECK(n_blocks < (1ull << 31), "cumulative column: too long");
    launch_fill_u32(c, d_flag, 1, 0);
    {
        KTimer kt(c, "logup_scan", (num_kind == 0 ? 16.0 : num_kind == 1 ? 4.0 : 0.0) * n + 32.0 * n);
        hipLaunchKernelGGL(logup_scan_tile_kernel, dim3((uint32_t)n_blocks), dim3(LG_THREADS), 0, c.stream, num_kind,
                           num, (const uint4*)den, n, (uint4*)sums, d_bs, d_flag);
        hipLaunchKernelGGL(logup_scan_blocks_kernel, dim3(1), dim3(LG_THREADS), 0, c.stream, d_bs, (uint32_t)n_blocks,
                           d_total);
        if (n_blocks > 1)
            hipLaunchKernelGGL(logup_scan_add_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, (uint4*)sums, n,
                               (const uint4*)d_bs);
        SR_HIP(hipGetLastError());
    }
}

qm31 logup_cumulative(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n, uint32_t* sums) {
    SR_CHECK(n > 0, "cumulative column: empty");
    DBuf tot(&c, 32), bs(&c, 16 * scan_blocks(n) + 16);
    uint4* d_total = tot.as<uint4>();
    launch_cumulative(c, num_kind, num, den, n, sums, bs.as<uint4>(), d_total, (uint32_t*)(d_total + 1));
    uint32_t* h = (uint32_t*)c.pinned(32);
    SR_HIP(hipMemcpyAsync(h, d_total, 20, hipMemcpyDeviceToHost, c.stream));
    c.sync();
    SR_CHECK(h[4] == 0, "LogUp: zero denominator");
    return qload(h);
}

void logup_cumulative_dev(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n, uint32_t* sums,
                          uint32_t* d_total, uint32_t* d_flag) {
    SR_CHECK(((uintptr_t)d_total & 15) == 0 && d_flag, "cumulative column: total must be 16-byte aligned");
    SR_CHECK(n > 0, "cumulative column: empty");
    DBuf bs(&c, 16 * scan_blocks(n) + 16);
    launch_cumulative(c, num_kind, num, den, n, sums, bs.as<uint4>(), (uint4*)d_total, d_flag);
}

// ---- the cumulative column without allocation ----
// A fixed per-device scratch array holds what the arena version allocates per call: the total
// (16 B) and the zero-denominator flag at offset 0, the block sums from offset 256.  Every call
// on a device shares that device's copy (the flag is cleared by a fill kernel each call; the
// total and the block sums are written before they are read), so calls that overlap on one GPU
// race.  The host never names the array itself: each device's address comes from
// hipGetSymbolAddress at context set-up.
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + 16 * (LOGUP_NOALLOC_MAX_ROWS / SCAN_TILE);
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[MAX_DEVICES] = {};

void logup_scratch_init(int device) {
    SR_CHECK(device >= 0 && device < MAX_DEVICES, "scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    SR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));  // the calling thread has `device` current
    SR_CHECK(p, "scratch: no device address");
    g_scratch_ptr[device] = (uint8_t*)p;
}

qm31 logup_cumulative_noalloc(Ctx& c, int num_kind, const uint32_t* num, const uint32_t* den, uint64_t n,
                              uint32_t* sums) {
    SR_CHECK(n > 0, "cumulative column: empty");
    SR_CHECK(n <= LOGUP_NOALLOC_MAX_ROWS, "cumulative column without allocation: at most 2^27 rows (the scratch "
                                          "array's size)");
    uint8_t* base;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = c.device >= 0 && c.device < MAX_DEVICES ? g_scratch_ptr[c.device] : nullptr;
    }
    SR_CHECK(base, "scratch: not set up for this device");
    uint4* d_total = (uint4*)base;
    launch_cumulative(c, num_kind, num, den, n, sums, (uint4*)(base + 256), d_total, (uint32_t*)(base + 16));
    uint32_t h[5];
    SR_HIP(hipMemcpyAsync(h, d_total, sizeof h, hipMemcpyDeviceToHost, c.stream));
    SR_HIP(hipStreamSynchronize(c.stream));
    SR_CHECK(h[4] == 0, "LogUp: zero denominator");
    return qload(h);
}

void launch_gkr_eq_evals(Ctx& c, const uint32_t* d_y, uint32_t n, qm31 scale, uint32_t* d_out) {
    SR_CHECK(n <= 30, "eq evals: too many variables");
    uint64_t len = 1ull << n;
    KTimer kt(c, "gkr_eq_evals", 16.0 * len);
    hipLaunchKernelGGL(gkr_eq_evals_kernel, grid_for(len), dim3(LG_THREADS), 0, c.stream, (const uint4*)d_y, n, scale,
                       (uint4*)d_out);
    SR_HIP(hipGetLastError());
}

void launch_gkr_next_layer(Ctx& c, int kind, const uint32_t* d_num, const uint32_t* d_den, uint32_t n_vars,
                           uint32_t* d_out_num, uint32_t* d_out_den) {
    SR_CHECK(kind >= 0 && kind <= 3, "gkr: layer kind");
    SR_CHECK(n_vars >= 1 && n_vars <= 30, "gkr next layer: variables");
    SR_CHECK(kind == GKR_GRAND_PRODUCT || kind == GKR_LOGUP_SINGLES || d_num, "gkr next layer: numerators");
    SR_CHECK(kind == GKR_GRAND_PRODUCT || d_out_num, "gkr next layer: output numerators");
    uint64_t half = 1ull << (n_vars - 1);
    double in_b = 32.0 + (kind == GKR_LOGUP_GENERIC ? 32.0 : kind == GKR_LOGUP_MULTIPLICITIES ? 8.0 : 0.0);
    KTimer kt(c, "gkr_next_layer", (in_b + (kind == GKR_GRAND_PRODUCT ? 16.0 : 32.0)) * half);
    hipLaunchKernelGGL(gkr_next_layer_kernel, grid_for(half), dim3(LG_THREADS), 0, c.stream, kind, d_num,
                       (const uint4*)d_den, half, (uint4*)d_out_num, (uint4*)d_out_den);
    SR_HIP(hipGetLastError());
}

void launch_gkr_sums(Ctx& c, int kind, const uint32_t* d_eq, const uint32_t* d_num, const uint32_t* d_den,
                     uint32_t n_vars, qm31 lambda, unsigned long long* d_acc8) {
    SR_CHECK(n_vars >= 2, "gkr sums: the oracle needs at least one variable");
    uint64_t nt = 1ull << (n_vars - 2);
    uint64_t blocks = (nt + LG_THREADS - 1) / LG_THREADS;
    if (blocks > 2048) blocks = 2048;
    KTimer kt(c, "gkr_sum_poly", (16.0 + 64.0 + (kind == GKR_LOGUP_GENERIC ? 64.0 : kind == GKR_LOGUP_MULTIPLICITIES ? 16.0
                                                                                                                  : 0.0)) *
                                     nt);
    hipLaunchKernelGGL(gkr_sum_kernel, dim3((uint32_t)blocks), dim3(LG_THREADS), 0, c.stream, kind,
                       (const uint4*)d_eq, d_num, (const uint4*)d_den, nt, lambda, d_acc8);
    SR_HIP(hipGetLastError());
}

void launch_mle_fix_first(Ctx& c, bool base, const uint32_t* d_in, uint32_t n_vars, qm31 r, uint32_t* d_out) {
    SR_CHECK(n_vars >= 1 && n_vars <= 30, "mle fix: variables");
    uint64_t half = 1ull << (n_vars - 1);
    KTimer kt(c, "mle_fix_first", ((base ? 8.0 : 32.0) + 16.0) * half);
    hipLaunchKernelGGL(mle_fix_first_kernel, grid_for(half), dim3(LG_THREADS), 0, c.stream, base ? 1 : 0, d_in, half,
                       r, (uint4*)d_out);
    SR_HIP(hipGetLastError());
}

void launch_logup_table_cols(Ctx& c, const uint32_t* mult16, uint64_t n, uint32_t* mpad, uint32_t* negm) {
    SR_CHECK(n >= (1u << 16), "logup table columns need at least 2^16 rows");
    KTimer kt(c, "logup_table_cols", 4.0 * (1u << 16) + 8.0 * n);
    hipLaunchKernelGGL(logup_table_cols_kernel, grid_for(n), dim3(LG_THREADS), 0, c.stream, mult16, n, mpad, negm);
    SR_HIP(hipGetLastError());
}

void launch_zero_u64(Ctx& c, unsigned long long* d, uint32_t n) {
    hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, d, n);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
