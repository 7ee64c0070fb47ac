// Circle FFT for gfx950: LDS-staged multi-stage passes over batched columns.
//
// Semantics (src/prover/poly/circle/transforms.zig:149-360, core/fft.zig:8-18):
// values are stored in bit-reversed circle-domain order.  A transform of log
// size m has stages s = m-1 .. 1 (line layers, butterfly half-block 2^s,
// twiddle T_{s-1}[h], h = idx0 >> (s+1)) and stage 0 (circle layer on adjacent
// pairs, twiddles [y, -y, -x, x] built from consecutive pairs (x, y) of T_0).
// Forward runs stages top-down with butterfly (v0 + v1*t, v0 - v1*t); inverse
// runs them bottom-up with (v0 + v1, (v0 - v1)*t) and scales by 2^-m.
//
// Twiddle tower: T_j of a domain of log m is the layer of length 2^(m-2-j) of
// CanonicCoset(M)'s tree (twiddle_tower.zig:66-87 suffix views); it starts at
// word 2^(M-1) - 2^(m-1-j) of the tower, so one HBM-resident tower serves
// every size.
//
// Pass structure: stages are split into passes of <= 11 stages.  The pass
// holding stage 0 works on contiguous 2^R-word chunks; higher passes work on
// tiles of 2^R strided rows x C consecutive words so that each row is a
// coalesced C-word segment.  Each element is read from and written to HBM
// once per pass.
#include "runtime.hpp"

namespace sr {

constexpr int FFT_THREADS = 256;
constexpr uint32_t FFT_MAX_TILE = 16384;  // words of LDS per block (64 KiB)

struct FftPass {
    uint32_t m, src_log;
    uint32_t s_lo, s_hi, log_c;
    uint32_t tower_log;
    uint32_t scale;  // applied on store when != 1
};

__device__ __forceinline__ uint32_t tw_at(const uint32_t* __restrict__ tower, uint32_t tower_log, uint32_t m,
                                          uint32_t s, uint32_t h) {
    // T_{s-1} of domain m starts at 2^(M-1) - 2^(m-s).
    return tower[((1u << (tower_log - 1)) - (1u << (m - s))) + h];
}

template <bool INVERSE>
__global__ __launch_bounds__(FFT_THREADS) void fft_pass_kernel(const uint32_t* const* __restrict__ src_cols,
                                                               uint32_t* const* __restrict__ dst_cols,
                                                               const uint32_t* __restrict__ tower, FftPass ps) {
    extern __shared__ uint32_t lds[];
    const uint32_t R = ps.s_hi - ps.s_lo;  // stages in this pass
    const uint32_t C = 1u << ps.log_c;
    const uint32_t tile = (1u << R) << ps.log_c;
    const uint32_t lo_blocks_log = ps.s_lo - ps.log_c;
    const uint32_t blocks_per_col = 1u << (ps.m - R - ps.log_c);
    const uint32_t col = blockIdx.x / blocks_per_col;
    const uint32_t b = blockIdx.x % blocks_per_col;
    const uint32_t lo_blk = b & ((1u << lo_blocks_log) - 1);
    const uint32_t hi = b >> lo_blocks_log;
    const uint32_t base = (hi << ps.s_hi) | (lo_blk << ps.log_c);
    const uint32_t* __restrict__ src = src_cols[col];
    uint32_t* __restrict__ dst = dst_cols[col];
    const uint32_t src_n = 1u << ps.src_log;

    const uint32_t stride = C;
    for (uint32_t e = threadIdx.x; e < tile; e += FFT_THREADS) {
        uint32_t j = e >> ps.log_c, lane = e & (C - 1);
        uint32_t idx = base | (j << ps.s_lo) | lane;
        lds[j * stride + lane] = idx < src_n ? src[idx] : 0u;
    }
    __syncthreads();

    const uint32_t n_bfly = tile >> 1;
    for (uint32_t k = 0; k < R; k++) {
        const uint32_t ls = INVERSE ? k : (R - 1 - k);
        const uint32_t s = ps.s_lo + ls;
        for (uint32_t t = threadIdx.x; t < n_bfly; t += FFT_THREADS) {
            uint32_t lane = t & (C - 1), pair = t >> ps.log_c;
            uint32_t j0 = ((pair >> ls) << (ls + 1)) | (pair & ((1u << ls) - 1));
            uint32_t j1 = j0 | (1u << ls);
            uint32_t tw;
            if (s > 0) {
                uint32_t h = (hi << (ps.s_hi - s - 1)) | (j0 >> (ls + 1));
                tw = tw_at(tower, ps.tower_log, ps.m, s, h);
            } else {
                // Circle stage: only in the contiguous pass (C == 1, s_lo == 0).
                uint32_t hp = (base | j0) >> 1;
                uint32_t kk = hp >> 2, r = hp & 3;
                const uint32_t* T0 = tower + ((1u << (ps.tower_log - 1)) - (1u << (ps.m - 1)));
                uint32_t v = (r == 0 || r == 1) ? T0[2 * kk + 1] : T0[2 * kk];
                tw = (r == 1 || r == 2) ? mneg(v) : v;
            }
            uint32_t* p0 = &lds[j0 * stride + lane];
            uint32_t* p1 = &lds[j1 * stride + lane];
            uint32_t v0 = *p0, v1 = *p1;
            if (!INVERSE) {
                uint32_t tmp = mmul(v1, tw);
                *p0 = madd(v0, tmp);
                *p1 = msub(v0, tmp);
            } else {
                *p0 = madd(v0, v1);
                *p1 = mmul(msub(v0, v1), tw);
            }
        }
        __syncthreads();
    }

    for (uint32_t e = threadIdx.x; e < tile; e += FFT_THREADS) {
        uint32_t j = e >> ps.log_c, lane = e & (C - 1);
        uint32_t idx = base | (j << ps.s_lo) | lane;
        uint32_t v = lds[j * stride + lane];
        if (ps.scale != 1) v = mmul(v, ps.scale);
        dst[idx] = v;
    }
}

// m <= 2: closed forms (transforms.zig:155-179,256-292), one thread per column.
__global__ void fft_small_kernel(const uint32_t* const* __restrict__ src_cols, uint32_t* const* __restrict__ dst_cols,
                                 uint32_t n_cols, uint32_t m, uint32_t src_log, bool inverse, uint32_t x, uint32_t y,
                                 uint32_t xi, uint32_t yi, uint32_t n_inv) {
    uint32_t col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols) return;
    const uint32_t* s = src_cols[col];
    uint32_t* d = dst_cols[col];
    uint32_t n = 1u << m, sn = 1u << src_log;
    uint32_t v[4];
    for (uint32_t i = 0; i < n; i++) v[i] = i < sn ? s[i] : 0u;
    auto bf = [](uint32_t& a, uint32_t& b, uint32_t t) {
        uint32_t tmp = mmul(b, t);
        b = msub(a, tmp);
        a = madd(a, tmp);
    };
    auto ibf = [](uint32_t& a, uint32_t& b, uint32_t t) {
        uint32_t tmp = a;
        a = madd(tmp, b);
        b = mmul(msub(tmp, b), t);
    };
    if (!inverse) {
        if (m == 1) {
            bf(v[0], v[1], y);
        } else {
            bf(v[0], v[2], x);
            bf(v[1], v[3], x);
            bf(v[0], v[1], y);
            bf(v[2], v[3], mneg(y));
        }
    } else {
        if (m == 1) {
            ibf(v[0], v[1], yi);
        } else {
            ibf(v[0], v[1], yi);
            ibf(v[2], v[3], mneg(yi));
            ibf(v[0], v[2], xi);
            ibf(v[1], v[3], xi);
        }
        for (uint32_t i = 0; i < n; i++) v[i] = mmul(v[i], n_inv);
    }
    for (uint32_t i = 0; i < n; i++) d[i] = v[i];
}

static void run_pass(Ctx& c, bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols,
                     const FftPass& ps) {
    uint32_t R = ps.s_hi - ps.s_lo;
    uint32_t tile = (1u << R) << ps.log_c;
    size_t lds = (size_t)tile * 4;
    uint64_t blocks = (uint64_t)n_cols << (ps.m - R - ps.log_c);
    SR_CHECK(blocks < (1ull << 31), "fft grid too large");
    SR_CHECK(tile >= 2, "fft tile too small");
    const uint32_t* tower = inverse ? c.d_itw : c.d_tw;
    if (inverse)
        hipLaunchKernelGGL(fft_pass_kernel<true>, dim3((uint32_t)blocks), dim3(FFT_THREADS), lds, c.stream, src, dst,
                           tower, ps);
    else
        hipLaunchKernelGGL(fft_pass_kernel<false>, dim3((uint32_t)blocks), dim3(FFT_THREADS), lds, c.stream, src,
                           dst, tower, ps);
    SR_HIP(hipGetLastError());
}

void launch_circle_fft(Ctx& c, bool inverse, const uint32_t* const* d_src, uint32_t* const* d_dst, uint32_t n_cols,
                       uint32_t m, uint32_t src_log) {
    if (n_cols == 0) return;
    SR_CHECK(m >= 1 && m <= 30, "circle fft: log size out of range");
    SR_CHECK(src_log <= m, "circle fft: source longer than domain");
    SR_CHECK(!inverse || src_log == m, "circle fft: inverse needs a full evaluation");
    if (m <= 2) {
        GenDoubles& g = c.gen;
        pt p0 = idx_to_point(g, half_odds_initial(m - 1));
        uint32_t n_inv = minv(1u << m);
        uint32_t xi = m == 2 ? minv(p0.x) : 0, yi = minv(p0.y);
        hipLaunchKernelGGL(fft_small_kernel, dim3((n_cols + 63) / 64), dim3(64), 0, c.stream, d_src, d_dst, n_cols, m,
                           src_log, inverse, p0.x, p0.y, xi, yi, n_inv);
        SR_HIP(hipGetLastError());
        return;
    }
    c.ensure_tower(m);
    // Pass plan: [0, low) contiguous, then strided passes of <= 10 stages.
    std::vector<std::pair<uint32_t, uint32_t>> passes;
    uint32_t low = m < 11 ? m : 11;
    passes.push_back({0, low});
    for (uint32_t s = low; s < m; s += 10) passes.push_back({s, s + 10 < m ? s + 10 : m});
    uint32_t n_inv = minv((uint32_t)((1ull << m) % P));
    auto mk = [&](std::pair<uint32_t, uint32_t> rg, bool last) {
        FftPass ps;
        ps.m = m;
        ps.s_lo = rg.first;
        ps.s_hi = rg.second;
        ps.tower_log = c.tower_log;
        uint32_t R = rg.second - rg.first;
        uint32_t log_c = 0;
        if (rg.first > 0) {
            // C consecutive words per strided row: >= 16 and a tile of ~4K..16K words.
            log_c = 4;
            while ((R + log_c) < 12 && log_c < rg.first) log_c++;
            if (log_c > rg.first) log_c = rg.first;
            while (((1u << R) << log_c) > FFT_MAX_TILE && log_c > 0) log_c--;
        }
        ps.log_c = log_c;
        ps.scale = (inverse && last) ? n_inv : 1u;
        return ps;
    };
    size_t np = passes.size();
    if (!inverse) {
        // top-down: last pass in the plan first; it reads the (short) source.
        for (size_t i = 0; i < np; i++) {
            size_t pi = np - 1 - i;
            FftPass ps = mk(passes[pi], false);
            ps.src_log = (i == 0) ? src_log : m;
            run_pass(c, false, i == 0 ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, ps);
        }
    } else {
        for (size_t i = 0; i < np; i++) {
            FftPass ps = mk(passes[i], i + 1 == np);
            ps.src_log = m;
            run_pass(c, true, i == 0 ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, ps);
        }
    }
}

// ---- tower and domain tables ----
// Layer of length 2^(k-1) holds the x-coordinates of the first half of
// Coset.halfOdds(k) in bit-reversed order (twiddles.zig:67-91); layers for
// k = M-1 .. 1 are concatenated and end with a trailing 1.
__global__ void tower_kernel(uint32_t M, uint32_t* __restrict__ tw, uint32_t* __restrict__ itw, GenDoubles g) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = 1ull << (M - 1);
    if (i >= total) return;
    if (i == total - 1) {
        tw[i] = 1;
        itw[i] = 1;
        return;
    }
    // find layer: offsets 2^(M-1) - 2^k .. 2^(M-1) - 2^(k-1) hold the layer of length 2^(k-1).
    uint64_t rem = total - i;  // in (2^(k-1), 2^k]
    uint32_t k = 64 - __builtin_clzll(rem - 1);
    uint64_t off = total - (1ull << k);
    uint32_t j = (uint32_t)(i - off);
    uint32_t lg = k - 1;
    uint32_t nat = bitrev32(j, lg);
    uint32_t init = half_odds_initial(k), step = subgroup_gen(k);
    pt p = idx_to_point(g, (uint32_t)((init + (uint64_t)step * nat) & ORDER_MASK));
    tw[i] = p.x;
    itw[i] = minv(p.x);
}

void launch_tower(Ctx& c, uint32_t log, uint32_t* d_tw, uint32_t* d_itw) {
    uint64_t n = 1ull << (log - 1);
    hipLaunchKernelGGL(tower_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream, log, d_tw, d_itw,
                       c.gen);
    SR_HIP(hipGetLastError());
}

// x[r], y[r] = CanonicCoset(log).circleDomain().at(bitrev(r)) (quotient_domain_walk.zig:1-25).
__global__ void domain_points_kernel(uint32_t log, uint32_t* __restrict__ xy, GenDoubles g) {
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 1u << log;
    if (r >= n) return;
    pt p = idx_to_point(g, domain_index_at(log, bitrev32(r, log)));
    xy[r] = p.x;
    xy[n + r] = p.y;
}
void launch_domain_points(Ctx& c, uint32_t log, uint32_t* d_xy) {
    uint32_t n = 1u << log;
    hipLaunchKernelGGL(domain_points_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, log, d_xy, c.gen);
    SR_HIP(hipGetLastError());
}

// inv_y[i] = 1 / halfOdds(log-1).at(bitrev(i)).y (folding.zig:316-335,558-570).
__global__ void circle_inv_y_kernel(uint32_t log, uint32_t* __restrict__ out, GenDoubles g) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 1u << (log - 1);
    if (i >= n) return;
    uint32_t init = half_odds_initial(log - 1), step = subgroup_gen(log - 1);
    pt p = idx_to_point(g, (uint32_t)((init + (uint64_t)step * bitrev32(i, log - 1)) & ORDER_MASK));
    out[i] = minv(p.y);
}
void launch_circle_inv_y(Ctx& c, uint32_t log, uint32_t* d_out) {
    uint32_t n = 1u << (log - 1);
    hipLaunchKernelGGL(circle_inv_y_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, log, d_out, c.gen);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
