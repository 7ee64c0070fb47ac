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
#include <algorithm>

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

// ---------------------------------------------------------------------------
// Register-radix pass (every transform of log size >= 6).
//
// A pass covers local stages [0, R) of a tile of 2^R rows x 2^LOGC lanes
// (8192 elements except for small transforms).  Each of the tile/16 threads
// owns 16 elements of one lane whose rows differ in 4 consecutive row bits
// [b, b+4); it runs up to 4 butterfly stages on them in registers (radix 16),
// then the tile is exchanged through LDS so that the next 4 row bits become
// thread-local: ceil(R/4) LDS exchanges per pass instead of R barriers.  R and
// LOGC are template parameters, so the row mapping, strides and twiddle
// offsets of every group are compile-time shifts.  Twiddles of one stage are
// consecutive tower words (one base + constant offsets).  The first group
// loads from HBM, the last one stores to HBM.
constexpr int RX_LOG = 4;  // 16 elements per thread
constexpr int RX = 1 << RX_LOG;

struct RxPass {
    uint32_t m, src_log;
    uint32_t s_lo;       // first stage of the pass
    uint32_t tower_log;
    uint32_t scale;      // multiply on store when != 1 (last inverse pass)
    uint32_t dup_top;    // forward: top stage pairs with zero padding (v1 = 0)
};

template <int LOGC>
__device__ __forceinline__ uint32_t rx_lds(uint32_t j, uint32_t lane) {
    // [row][lane]; contiguous tiles are skewed by one word per 16 rows.
    return LOGC == 0 ? (j + (j >> 4)) : ((j << LOGC) | lane);
}

// LDS words of a thread's 16 values with row bits [B, B+4) (row0 has those bits zero): for
// interleaved columns, for B == 0 and for B >= 4 the skewed index of value k is the index of
// value 0 plus k * STRIDE (the row bits of k never carry into the skew), so the 16 accesses
// take one address and immediate offsets; B = 1..3 of a contiguous tile index each value.
template <int LOGC, int B>
struct RxLdsMap {
    static constexpr bool AFFINE = LOGC != 0 || B == 0 || B >= RX_LOG;
    static constexpr uint32_t STRIDE =
        LOGC != 0 ? (1u << (B + LOGC)) : (B == 0 ? 1u : (1u << B) + (1u << (B >= RX_LOG ? B - RX_LOG : 0)));
};

template <int LOGC, int B>
__device__ __forceinline__ void rx_lds_put(uint32_t* lds, const uint32_t (&v)[RX], uint32_t row0, uint32_t lane) {
    using M = RxLdsMap<LOGC, B>;
    if constexpr (M::AFFINE) {
        uint32_t* p = lds + rx_lds<LOGC>(row0, lane);
#pragma unroll
        for (int k = 0; k < RX; k++) p[k * M::STRIDE] = v[k];
    } else {
#pragma unroll
        for (int k = 0; k < RX; k++) lds[rx_lds<LOGC>(row0 + ((uint32_t)k << B), lane)] = v[k];
    }
}

template <int LOGC, int B>
__device__ __forceinline__ void rx_lds_get(uint32_t (&v)[RX], const uint32_t* lds, uint32_t row0, uint32_t lane) {
    using M = RxLdsMap<LOGC, B>;
    if constexpr (M::AFFINE) {
        const uint32_t* p = lds + rx_lds<LOGC>(row0, lane);
#pragma unroll
        for (int k = 0; k < RX; k++) v[k] = p[k * M::STRIDE];
    } else {
#pragma unroll
        for (int k = 0; k < RX; k++) v[k] = lds[rx_lds<LOGC>(row0 + ((uint32_t)k << B), lane)];
    }
}

// Twiddles of one stage (doubled tower words): RX >> (Q+1) consecutive words of T_{s-1}
// for a line stage; for the circle stage the two (x, y) pairs of T_0 the thread needs.
template <bool INV, int R, int LOGC, int B, int Q>
__device__ __forceinline__ void rx_tw_load(uint32_t (&w)[8], const uint32_t* __restrict__ tower, const RxPass& ps,
                                           uint32_t hi, uint32_t r, uint32_t base) {
    constexpr int LS = B + Q;
    // a contiguous pass (LOGC == 0) starts at stage 0 (the launcher checks it), so its local
    // stage 0 is the circle stage; interleaved passes start above it
    constexpr bool IS_CIRCLE = (LS == 0 && LOGC == 0);
    const uint32_t s = ps.s_lo + LS;
    if constexpr (!IS_CIRCLE) {
        // h = (hi << (R-LS-1)) | (r_high << (3-Q)) | hk  with r_high = r >> B
        const uint32_t tb = ((1u << (ps.tower_log - 1)) - (1u << (ps.m - s))) + ((hi << (R - LS - 1)) |
                                                                                 ((r >> B) << (3 - Q)));
        const uint32_t* __restrict__ twp = tower + tb;  // constant offsets below
#pragma unroll
        for (int hk = 0; hk < (RX >> (Q + 1)); hk++) w[hk] = twp[hk];
    } else {
        const uint32_t* T0 = tower + ((1u << (ps.tower_log - 1)) - (1u << (ps.m - 1)));
        const uint32_t kk0 = ((base >> 1) | (r << 3)) >> 2;
#pragma unroll
        for (int k = 0; k < 4; k++) w[k] = T0[2 * kk0 + k];
    }
}

template <bool INV, int R, int LOGC, int B, int Q>
__device__ __forceinline__ void rx_stage(uint32_t (&v)[RX], const uint32_t (&w)[8], const RxPass& ps, bool dup) {
    constexpr int LS = B + Q;
    constexpr bool IS_CIRCLE = (LS == 0 && LOGC == 0);  // as in rx_tw_load
    if constexpr (!IS_CIRCLE) {
#pragma unroll
        for (int hk = 0; hk < (RX >> (Q + 1)); hk++) {
            const uint32_t tw = w[hk];
#pragma unroll
            for (int lk = 0; lk < (1 << Q); lk++) {
                const int k0 = (hk << (Q + 1)) | lk, k1 = k0 | (1 << Q);
                uint32_t a = v[k0], c = v[k1];
                if (!INV) {
                    if (dup) {
                        v[k1] = a;  // zero-padded partner: (a + 0*t, a - 0*t)
                    } else {
                        uint32_t tmp = mmul_t2(c, tw);
                        v[k0] = madd(a, tmp);
                        v[k1] = msub(a, tmp);
                    }
                } else {
                    v[k0] = madd(a, c);
                    v[k1] = mmul_t2(msub(a, c), tw);
                }
            }
        }
    } else {
        // Circle stage (B == 0, Q == 0, s_lo == 0): pairs (2h', 2h'+1), h' = hp0 + hk,
        // twiddles [y, -y, -x, x] from consecutive pairs (x, y) of T_0.
        const uint32_t x0 = w[0], y0 = w[1], x1 = w[2], y1 = w[3];
#pragma unroll
        for (int hk = 0; hk < RX / 2; hk++) {
            const int k0 = 2 * hk, k1 = k0 + 1;
            const uint32_t x = hk < 4 ? x0 : x1, y = hk < 4 ? y0 : y1;
            const int rr = hk & 3;
            const uint32_t wv = rr < 2 ? y : x;
            const uint32_t t = (rr == 1 || rr == 2) ? mneg_t2(wv) : wv;  // doubled words
            uint32_t a = v[k0], c = v[k1];
            if (!INV) {
                uint32_t tmp = mmul_t2(c, t);
                v[k0] = madd(a, tmp);
                v[k1] = msub(a, tmp);
            } else {
                v[k0] = madd(a, c);
                v[k1] = mmul_t2(msub(a, c), t);
            }
        }
    }
}

// One group: row bits [B, B+4) owned, active local stages [LO, HI); twiddles w[Q] of
// stage B + Q are loaded ahead (rx_group_tw).
template <bool INV, int R, int LOGC, int B, int LO, int HI>
__device__ __forceinline__ void rx_group_tw(uint32_t (&w)[4][8], const uint32_t* __restrict__ tower, const RxPass& ps,
                                            uint32_t hi, uint32_t r, uint32_t base) {
    if (B + 3 >= LO && B + 3 < HI) rx_tw_load<INV, R, LOGC, B, 3>(w[3], tower, ps, hi, r, base);
    if (B + 2 >= LO && B + 2 < HI) rx_tw_load<INV, R, LOGC, B, 2>(w[2], tower, ps, hi, r, base);
    if (B + 1 >= LO && B + 1 < HI) rx_tw_load<INV, R, LOGC, B, 1>(w[1], tower, ps, hi, r, base);
    if (B + 0 >= LO && B + 0 < HI) rx_tw_load<INV, R, LOGC, B, 0>(w[0], tower, ps, hi, r, base);
}

template <bool INV, int R, int LOGC, int B, int LO, int HI>
__device__ __forceinline__ void rx_group(uint32_t (&v)[RX], const uint32_t (&w)[4][8], const RxPass& ps, bool dup) {
    if (!INV) {
        if (B + 3 >= LO && B + 3 < HI) rx_stage<INV, R, LOGC, B, 3>(v, w[3], ps, dup);
        if (B + 2 >= LO && B + 2 < HI) rx_stage<INV, R, LOGC, B, 2>(v, w[2], ps, false);
        if (B + 1 >= LO && B + 1 < HI) rx_stage<INV, R, LOGC, B, 1>(v, w[1], ps, false);
        if (B + 0 >= LO && B + 0 < HI) rx_stage<INV, R, LOGC, B, 0>(v, w[0], ps, false);
    } else {
        if (B + 0 >= LO && B + 0 < HI) rx_stage<INV, R, LOGC, B, 0>(v, w[0], ps, false);
        if (B + 1 >= LO && B + 1 < HI) rx_stage<INV, R, LOGC, B, 1>(v, w[1], ps, false);
        if (B + 2 >= LO && B + 2 < HI) rx_stage<INV, R, LOGC, B, 2>(v, w[2], ps, false);
        if (B + 3 >= LO && B + 3 < HI) rx_stage<INV, R, LOGC, B, 3>(v, w[3], ps, false);
    }
}

// Group gi (compile time): inverse bottom-up, forward top-down over the same row windows
// [B, B+4) with B = 0, 4, 8, ... and the topmost window aligned to the tile edge (B = R - 4,
// fewer active stages when 4 does not divide R).  Every B is then 0 or >= 4 once R >= 8,
// which keeps the LDS exchanges affine (RxLdsMap).
template <bool INV, int R, int GI>
struct RxGroupShape {
    static constexpr int NG = (R + RX_LOG - 1) / RX_LOG;
    static constexpr int G = INV ? GI : NG - 1 - GI;  // window index from the bottom
    static constexpr int LO = G * RX_LOG;
    static constexpr int HI = G * RX_LOG + RX_LOG < R ? G * RX_LOG + RX_LOG : R;
    static constexpr int B = G * RX_LOG + RX_LOG <= R ? G * RX_LOG : R - RX_LOG;
};

// w holds this group's twiddles (loaded before the previous exchange); the next group's
// twiddles are loaded before this group's exchange so their latency overlaps it.
// A row's uniform address kept in SGPRs and hidden from reassociation (an empty asm, no
// instruction).  Otherwise the compiler adds the lane offset to the column pointer once, in a
// 64-bit VGPR base, and pays one 64-bit VALU add per access for the row offsets; this way
// every strided access takes the SGPR-base + 32-bit lane offset form.
template <class T>
__device__ __forceinline__ T* rx_sbase(T* p) {
    uint64_t v = (uint64_t)p;
    asm("" : "+s"(v));
    return (T*)v;
}

// FROM_REGS: the first group takes its 16 values from v instead of loading them (the fused
// interpolate + extend kernel hands the coefficients over in registers).
// SP (compile time): an inverse pass scales on store (ps.scale), a forward pass has a
// zero-padded top stage (ps.dup_top): its values 8..15 are the zero partners, never read.
// Strided HBM accesses address the uniform column pointer plus a uniform row offset (rx_sbase)
// with a 32-bit per-lane byte offset (words < 2^30), which the SGPR-base global forms take as is.
template <bool INV, int R, int LOGC, int GI, bool SP, bool FROM_REGS = false>
__device__ __forceinline__ void rx_run(uint32_t (&v)[RX], const uint32_t (&w)[4][8], uint32_t* lds,
                                       gcu32* __restrict__ src, gu32* __restrict__ dst,
                                       const uint32_t* __restrict__ tower, const RxPass& ps, uint32_t hi, uint32_t r,
                                       uint32_t lane, uint32_t base) {
    constexpr int NG = (R + RX_LOG - 1) / RX_LOG;
    constexpr bool DUP = !INV && SP;
    if constexpr (GI < NG) {
        using S = RxGroupShape<INV, R, GI>;
        constexpr int B = S::B;
        const uint32_t row0 = (r & ((1u << B) - 1)) | ((r >> B) << (B + RX_LOG));
        if (GI == 0 && FROM_REGS) {
            // values already in v (row bits [B, B+4) of this thread)
        } else if (GI == 0) {
            const uint32_t g0 = base | (row0 << ps.s_lo) | lane;
            const uint32_t src_n = 1u << ps.src_log;
            if (B == 0 && LOGC == 0 && g0 + RX <= src_n) {
                // the thread's 16 words are consecutive (g0 is a multiple of 16): 16-byte loads
                const SR_GLOBAL u32x4* s4 = (const SR_GLOBAL u32x4*)(src + g0);
#pragma unroll
                for (int k = 0; k < RX / 4; k++) {
                    const u32x4 w = s4[k];
                    v[4 * k] = w.x;
                    v[4 * k + 1] = w.y;
                    v[4 * k + 2] = w.z;
                    v[4 * k + 3] = w.w;
                }
            } else if (INV || ps.src_log >= ps.m) {
                // a full source (every inverse pass, every forward pass after the first):
                // no row is past its end, so the loads take no bound check
                const uint32_t sh = B + ps.s_lo, voff = g0 << 2;
#pragma unroll
                for (int k = 0; k < RX; k++) {
                    const SR_GLOBAL uint8_t* sk = rx_sbase((const SR_GLOBAL uint8_t*)src + ((uint64_t)k << (sh + 2)));
                    v[k] = *(gcu32*)(sk + voff);
                }
            } else {
                const uint32_t sh = B + ps.s_lo, voff = g0 << 2;
#pragma unroll
                for (int k = 0; k < RX; k++) {
                    if (DUP && k >= RX / 2) {
                        v[k] = 0u;
                        continue;
                    }
                    const SR_GLOBAL uint8_t* sk = (const SR_GLOBAL uint8_t*)src + ((uint64_t)k << (sh + 2));
                    v[k] = g0 + ((uint32_t)k << sh) < src_n ? *(gcu32*)(sk + voff) : 0u;
                }
            }
        } else {
            __syncthreads();
            rx_lds_get<LOGC, B>(v, lds, row0, lane);
        }
        rx_group<INV, R, LOGC, B, S::LO, S::HI>(v, w, ps, DUP && GI == 0);
        if constexpr (GI + 1 < NG) {
            using SN = RxGroupShape<INV, R, GI + 1>;
            uint32_t wn[4][8];
            rx_group_tw<INV, R, LOGC, SN::B, SN::LO, SN::HI>(wn, tower, ps, hi, r, base);
            if (GI > 0) __syncthreads();
            rx_lds_put<LOGC, B>(lds, v, row0, lane);
            rx_run<INV, R, LOGC, GI + 1, SP, FROM_REGS>(v, wn, lds, src, dst, tower, ps, hi, r, lane, base);
        } else {
            const uint32_t g0 = base | (row0 << ps.s_lo) | lane;
            if constexpr (INV && SP) {
#pragma unroll
                for (int k = 0; k < RX; k++) v[k] = mmul(v[k], ps.scale);
            }
            if constexpr (B == 0 && LOGC == 0) {
                SR_GLOBAL u32x4* d4 = (SR_GLOBAL u32x4*)(dst + g0);
#pragma unroll
                for (int k = 0; k < RX / 4; k++) d4[k] = u32x4{v[4 * k], v[4 * k + 1], v[4 * k + 2], v[4 * k + 3]};
            } else {
                const uint32_t sh = B + ps.s_lo, voff = g0 << 2;
#pragma unroll
                for (int k = 0; k < RX; k++)
                    *(gu32*)(rx_sbase((SR_GLOBAL uint8_t*)dst + ((uint64_t)k << (sh + 2))) + voff) = v[k];
            }
        }
    }
}

template <bool INV, int R, int LOGC, bool SP>
__global__ __launch_bounds__(512) void fft_rx_kernel(const uint32_t* const* __restrict__ src_cols,
                                                     uint32_t* const* __restrict__ dst_cols,
                                                     const uint32_t* __restrict__ tower, RxPass ps) {
    extern __shared__ uint32_t lds[];
    constexpr int TILE_LOG = R + LOGC;
    const uint32_t per_col_log = ps.m - TILE_LOG;
    const uint32_t col = blockIdx.x >> per_col_log;
    const uint32_t bb = blockIdx.x & ((1u << per_col_log) - 1);
    const uint32_t lo_blocks_log = ps.s_lo - LOGC;
    const uint32_t lo_blk = bb & ((1u << lo_blocks_log) - 1);
    const uint32_t hi = bb >> lo_blocks_log;
    const uint32_t base = (hi << (ps.s_lo + R)) | (lo_blk << LOGC);
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & ((1u << LOGC) - 1);
    const uint32_t r = t >> LOGC;
    uint32_t v[RX];
    uint32_t w[4][8];
    using S0 = RxGroupShape<INV, R, 0>;
    rx_group_tw<INV, R, LOGC, S0::B, S0::LO, S0::HI>(w, tower, ps, hi, r, base);
    rx_run<INV, R, LOGC, 0, SP>(v, w, lds, (gcu32*)src_cols[col], (gu32*)dst_cols[col], tower, ps, hi, r, lane, base);
}

// Interpolation of log n fused with the extension to log n + 1 (blowup 2), over the top
// strided pass of both transforms (transforms.zig:249-360 then 149-246).  The inverse pass
// (stages [s_lo, n) of the evaluation domain, scaled by 2^-n) leaves each thread's 16
// coefficients in registers; they are stored (the retained coefficients) and extended: the
// top forward stage n pairs every coefficient with a zero, so both halves of the extended
// column start from the same coefficients, and each half runs the forward stages
// [s_lo, n) of CanonicCoset(n + 1) with its own twiddles (hi = the half).  The inverse's last
// row group and the forward's first own the same row bits, so no exchange separates them.
// Saves one HBM read of the coefficients and one launch per column set.
template <int R, int LOGC>
__global__ __launch_bounds__(512) void fft_ifft_lde_kernel(const uint32_t* const* __restrict__ src_cols,
                                                           uint32_t* const* __restrict__ coef_cols,
                                                           uint32_t* const* __restrict__ lde_cols,
                                                           const uint32_t* __restrict__ itower,
                                                           const uint32_t* __restrict__ tower, RxPass pi, RxPass pf) {
    extern __shared__ uint32_t lds[];
    constexpr int TILE_LOG = R + LOGC;
    static_assert(RxGroupShape<true, R, (R + RX_LOG - 1) / RX_LOG - 1>::B == RxGroupShape<false, R, 0>::B,
                  "the inverse's last group and the forward's first must own the same row bits");
    const uint32_t per_col_log = pi.m - TILE_LOG;  // pi.s_lo + R == pi.m: every block bit is a lane block
    const uint32_t col = blockIdx.x >> per_col_log;
    const uint32_t lo_blk = blockIdx.x & ((1u << per_col_log) - 1);
    const uint32_t base = lo_blk << LOGC;
    const uint32_t t = threadIdx.x;
    const uint32_t lane = t & ((1u << LOGC) - 1);
    const uint32_t r = t >> LOGC;
    uint32_t v[RX];
    uint32_t w[4][8];
    {
        using S0 = RxGroupShape<true, R, 0>;
        rx_group_tw<true, R, LOGC, S0::B, S0::LO, S0::HI>(w, itower, pi, 0, r, base);
        rx_run<true, R, LOGC, 0, true>(v, w, lds, (gcu32*)src_cols[col], (gu32*)coef_cols[col], itower, pi, 0, r, lane,
                                 base);
    }
    uint32_t cf[RX];
#pragma unroll
    for (int k = 0; k < RX; k++) cf[k] = v[k];
    gu32* lde = (gu32*)lde_cols[col];
    using F0 = RxGroupShape<false, R, 0>;
#pragma unroll
    for (uint32_t h = 0; h < 2; h++) {
        if (h) {
#pragma unroll
            for (int k = 0; k < RX; k++) v[k] = cf[k];
        }
        rx_group_tw<false, R, LOGC, F0::B, F0::LO, F0::HI>(w, tower, pf, h, r, base | (h << (pf.s_lo + R)));
        __syncthreads();  // the previous transform's last exchange has been read
        rx_run<false, R, LOGC, 0, false, true>(v, w, lds, nullptr, lde, tower, pf, h, r, lane,
                                               base | (h << (pf.s_lo + R)));
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
    KTimer kt(c, inverse ? "fft_ifft_pass" : "fft_lde_pass",
              4.0 * n_cols * ((double)(1ull << (ps.src_log < ps.m ? ps.src_log : ps.m)) + (double)(1ull << ps.m)));
    if (inverse)
        hipLaunchKernelGGL(fft_pass_kernel<true>, dim3((uint32_t)blocks), dim3(FFT_THREADS), lds, c.stream, src, dst,
                           tower, ps);
    else
        hipLaunchKernelGGL(fft_pass_kernel<false>, dim3((uint32_t)blocks), dim3(FFT_THREADS), lds, c.stream, src,
                           dst, tower, ps);
    SR_HIP(hipGetLastError());
}

template <bool INV>
static void rx_dispatch(uint32_t R, uint32_t logc, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                        const uint32_t* const* src, uint32_t* const* dst, const uint32_t* tower, const RxPass& ps) {
    // the inverse's scaled pass and the forward's zero-padded top pass run the SP form
    const bool sp = INV ? ps.scale != 1u : ps.dup_top != 0u;
#define SR_RX(RR, CC)                                                                                       \
    if (R == RR && logc == CC) {                                                                            \
        if (sp)                                                                                             \
            hipLaunchKernelGGL((fft_rx_kernel<INV, RR, CC, true>), grid, block, lds, st, src, dst, tower, ps); \
        else                                                                                                \
            hipLaunchKernelGGL((fft_rx_kernel<INV, RR, CC, false>), grid, block, lds, st, src, dst, tower, ps); \
        return;                                                                                             \
    }
    SR_RX(6, 0) SR_RX(7, 0) SR_RX(8, 0) SR_RX(9, 0) SR_RX(10, 0) SR_RX(11, 0) SR_RX(12, 0) SR_RX(13, 0)
    SR_RX(4, 9) SR_RX(5, 8) SR_RX(6, 7) SR_RX(7, 6) SR_RX(8, 5)  // 8192-word tiles
    SR_RX(4, 8) SR_RX(5, 7) SR_RX(6, 6) SR_RX(7, 5) SR_RX(8, 4)  // 4096-word tiles
    SR_RX(4, 7) SR_RX(5, 6) SR_RX(6, 5) SR_RX(7, 4) SR_RX(8, 3)  // 2048-word tiles
#undef SR_RX
    throw Error(1, "fft: no kernel for pass shape R=" + std::to_string(R) + " logc=" + std::to_string(logc));
}

static void run_rx_pass(Ctx& c, bool inverse, const uint32_t* const* src, uint32_t* const* dst, uint32_t n_cols,
                        uint32_t R, uint32_t logc, const RxPass& ps) {
    uint32_t tile_log = R + logc;
    uint32_t nt = 1u << (tile_log - RX_LOG);
    size_t lds = logc == 0 ? ((size_t)(1u << tile_log) + ((1u << tile_log) >> 4)) * 4 : ((size_t)4 << tile_log);
    uint64_t blocks = (uint64_t)n_cols << (ps.m - tile_log);
    SR_CHECK(blocks < (1ull << 31), "fft grid too large");
    SR_CHECK(R >= (uint32_t)RX_LOG && nt <= 512 && logc <= ps.s_lo && tile_log <= ps.m, "fft pass shape");
    SR_CHECK(logc != 0 || ps.s_lo == 0, "fft: contiguous pass must start at stage 0");
    const uint32_t* tower = inverse ? c.d_itw2 : c.d_tw2;  // doubled: products via mmul_t2
    KTimer kt(c,
              std::string(inverse ? "fft_ifft_pass" : "fft_lde_pass") + " m=" + std::to_string(ps.m) +
                  " R=" + std::to_string(R) + " C=" + std::to_string(1u << logc) + " cols=" + std::to_string(n_cols),
              4.0 * n_cols * ((double)(1ull << (ps.src_log < ps.m ? ps.src_log : ps.m)) + (double)(1ull << ps.m)));
    if (inverse)
        rx_dispatch<true>(R, logc, dim3((uint32_t)blocks), dim3(nt), lds, c.stream, src, dst, tower, ps);
    else
        rx_dispatch<false>(R, logc, dim3((uint32_t)blocks), dim3(nt), lds, c.stream, src, dst, tower, ps);
    SR_HIP(hipGetLastError());
}

// Pass plan of a transform of log m >= 6: a contiguous pass of R0 <= TILE_LOG stages, then
// strided passes of 4..8 stages over 2^TILE_LOG-element tiles (>= 8 consecutive words per row).
static std::vector<std::pair<uint32_t, uint32_t>> rx_plan(uint32_t m, uint32_t TILE_LOG) {
    std::vector<std::pair<uint32_t, uint32_t>> passes;  // (s_lo, R)
    if (m <= TILE_LOG) {
        passes.push_back({0, m});
    } else {
        uint32_t n_high = (m - TILE_LOG + 7) / 8;
        uint32_t r0 = std::min<uint32_t>(TILE_LOG, m - 4 * n_high);
        passes.push_back({0, r0});
        uint32_t rem = m - r0, s = r0;
        for (uint32_t i = 0; i < n_high; i++) {
            uint32_t left = n_high - i;
            uint32_t rr = (rem + left - 1) / left;
            passes.push_back({s, rr});
            s += rr;
            rem -= rr;
        }
    }
    return passes;
}
static uint32_t rx_logc(std::pair<uint32_t, uint32_t> p, uint32_t TILE_LOG) {
    return p.first == 0 ? 0u : std::min<uint32_t>(TILE_LOG - p.second, p.first);
}

static bool fused_dispatch(uint32_t R, uint32_t logc, dim3 grid, dim3 block, size_t lds, hipStream_t st,
                           const uint32_t* const* src, uint32_t* const* coef, uint32_t* const* lde,
                           const uint32_t* itower, const uint32_t* tower, const RxPass& pi, const RxPass& pf) {
#define SR_FX(RR, CC)                                                                                          \
    if (R == RR && logc == CC) {                                                                               \
        hipLaunchKernelGGL((fft_ifft_lde_kernel<RR, CC>), grid, block, lds, st, src, coef, lde, itower, tower, \
                           pi, pf);                                                                            \
        return true;                                                                                           \
    }
    SR_FX(4, 9) SR_FX(5, 8) SR_FX(6, 7) SR_FX(7, 6) SR_FX(8, 5)  // 8192-word tiles
    SR_FX(4, 8) SR_FX(5, 7) SR_FX(6, 6) SR_FX(7, 5) SR_FX(8, 4)  // 4096-word tiles
    SR_FX(4, 7) SR_FX(5, 6) SR_FX(6, 5) SR_FX(7, 4) SR_FX(8, 3)  // 2048-word tiles
#undef SR_FX
    return false;
}

void launch_circle_ifft_lde(Ctx& c, const uint32_t* const* d_eval, uint32_t* const* d_coef, uint32_t* const* d_lde,
                            uint32_t n_cols, uint32_t n) {
    if (n_cols == 0) return;
    const uint32_t TILE_LOG = c.fft_tile_log;
    const uint32_t m = n + 1;
    bool fuse = n > TILE_LOG && n >= 6;
    std::vector<std::pair<uint32_t, uint32_t>> pi_plan, pf_plan;
    if (fuse) {
        pi_plan = rx_plan(n, TILE_LOG);
        pf_plan = rx_plan(m, TILE_LOG);
        // the inverse's top pass and the extension's top pass cover the same stages [S, n) (the
        // extension adds stage n, its zero-padded duplication) on the same tile shape
        const auto a = pi_plan.back(), b = pf_plan.back();
        fuse = a.first == b.first && a.second + 1 == b.second && pi_plan.size() == pf_plan.size() &&
               a.first + a.second == n;
        for (size_t i = 0; fuse && i + 1 < pi_plan.size(); i++) fuse = pi_plan[i] == pf_plan[i];
    }
    if (!fuse) {
        launch_circle_fft(c, true, d_eval, d_coef, n_cols, n, n);
        launch_circle_fft(c, false, (const uint32_t* const*)d_coef, d_lde, n_cols, m, n);
        return;
    }
    c.ensure_tower(m);
    const uint32_t n_inv = minv((uint32_t)((1ull << n) % P));
    const size_t np = pi_plan.size();
    // inverse passes below the top one (evaluations -> d_coef, in place afterwards)
    for (size_t i = 0; i + 1 < np; i++) {
        RxPass ps{};
        ps.m = n;
        ps.src_log = n;
        ps.s_lo = pi_plan[i].first;
        ps.tower_log = c.tower_log;
        ps.scale = 1u;
        ps.dup_top = 0;
        run_rx_pass(c, true, i == 0 ? d_eval : (const uint32_t* const*)d_coef, d_coef, n_cols, pi_plan[i].second,
                    rx_logc(pi_plan[i], TILE_LOG), ps);
    }
    {
        const auto top = pi_plan.back();
        const uint32_t R = top.second, logc = rx_logc(top, TILE_LOG);
        RxPass pi{}, pf{};
        pi.m = n;
        pi.src_log = n;
        pi.s_lo = top.first;
        pi.tower_log = c.tower_log;
        pi.scale = n_inv;
        pf.m = m;
        pf.src_log = m;
        pf.s_lo = top.first;
        pf.tower_log = c.tower_log;
        pf.scale = 1u;
        const uint32_t tile_log = R + logc;
        SR_CHECK(R >= (uint32_t)RX_LOG && tile_log <= n && logc <= top.first, "fused fft pass shape");
        const uint32_t nt = 1u << (tile_log - RX_LOG);
        const size_t lds = (size_t)4 << tile_log;
        const uint64_t blocks = (uint64_t)n_cols << (n - tile_log);
        SR_CHECK(blocks < (1ull << 31), "fft grid too large");
        // in the middle of a plan with more than one inverse pass, the top pass reads d_coef
        const uint32_t* const* src = np > 1 ? (const uint32_t* const*)d_coef : d_eval;
        KTimer kt(c,
                  "fft_ilde_pass m=" + std::to_string(n) + " R=" + std::to_string(R) + " C=" +
                      std::to_string(1u << logc) + " cols=" + std::to_string(n_cols),
                  4.0 * n_cols * (2.0 * (double)(1ull << n) + (double)(1ull << m)));
        SR_CHECK(fused_dispatch(R, logc, dim3((uint32_t)blocks), dim3(nt), lds, c.stream, src, d_coef, d_lde,
                                c.d_itw2, c.d_tw2, pi, pf),
                 "fft: no fused kernel for the pass shape");
        SR_HIP(hipGetLastError());
    }
    // the extension's passes below the top one, in place on d_lde
    for (size_t i = 1; i < np; i++) {
        const auto pp = pf_plan[np - 1 - i];
        RxPass ps{};
        ps.m = m;
        ps.src_log = m;
        ps.s_lo = pp.first;
        ps.tower_log = c.tower_log;
        ps.scale = 1u;
        ps.dup_top = 0;
        run_rx_pass(c, false, (const uint32_t* const*)d_lde, d_lde, n_cols, pp.second, rx_logc(pp, TILE_LOG), ps);
    }
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
    uint32_t n_inv = minv((uint32_t)((1ull << m) % P));
    if (m < 6) {
        // Tiny transforms: one LDS pass, one stage per barrier.
        FftPass ps;
        ps.m = m;
        ps.src_log = src_log;
        ps.s_lo = 0;
        ps.s_hi = m;
        ps.log_c = 0;
        ps.tower_log = c.tower_log;
        ps.scale = inverse ? n_inv : 1u;
        run_pass(c, inverse, d_src, d_dst, n_cols, ps);
        return;
    }
    const uint32_t TILE_LOG = c.fft_tile_log;  // 11..13 (tuning parameter, results independent of it)
    const std::vector<std::pair<uint32_t, uint32_t>> passes = rx_plan(m, TILE_LOG);  // (s_lo, R)
    auto logc_of = [&](std::pair<uint32_t, uint32_t> p) { return rx_logc(p, TILE_LOG); };
    auto mk = [&](std::pair<uint32_t, uint32_t> p, bool last) {
        RxPass ps;
        ps.m = m;
        ps.s_lo = p.first;
        ps.tower_log = c.tower_log;
        ps.scale = (inverse && last) ? n_inv : 1u;
        ps.dup_top = 0;
        return ps;
    };
    size_t np = passes.size();
    if (!inverse) {
        for (size_t i = 0; i < np; i++) {
            auto pp = passes[np - 1 - i];
            RxPass ps = mk(pp, false);
            ps.src_log = i == 0 ? src_log : m;
            // The top stage pairs row i with the zero-padded row i + 2^(m-1).
            ps.dup_top = (i == 0 && src_log < m) ? 1u : 0u;
            run_rx_pass(c, false, i == 0 ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, pp.second,
                        logc_of(pp), ps);
        }
    } else {
        for (size_t i = 0; i < np; i++) {
            RxPass ps = mk(passes[i], i + 1 == np);
            ps.src_log = m;
            run_rx_pass(c, true, i == 0 ? d_src : (const uint32_t* const*)d_dst, d_dst, n_cols, passes[i].second,
                        logc_of(passes[i]), ps);
        }
    }
}

// ---- tower and domain tables ----
// Layer of length 2^(k-1) holds the x-coordinates of the first half of
// Coset.halfOdds(k) in bit-reversed order (twiddles.zig:67-91); layers for
// k = M-1 .. 1 are concatenated and end with a trailing 1.
__global__ void tower_kernel(uint32_t M, uint32_t* __restrict__ tw, uint32_t* __restrict__ itw,
                             uint32_t* __restrict__ tw2, uint32_t* __restrict__ itw2, GenDoubles g) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t total = 1ull << (M - 1);
    if (i >= total) return;
    if (i == total - 1) {
        tw[i] = 1;
        itw[i] = 1;
        tw2[i] = 2;
        itw2[i] = 2;
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
    const uint32_t ix = minv(p.x);
    tw[i] = p.x;
    itw[i] = ix;
    tw2[i] = p.x << 1;
    itw2[i] = ix << 1;
}

void launch_tower(Ctx& c, uint32_t log, uint32_t* d_tw, uint32_t* d_itw, uint32_t* d_tw2, uint32_t* d_itw2) {
    uint64_t n = 1ull << (log - 1);
    hipLaunchKernelGGL(tower_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, c.stream, log, d_tw, d_itw,
                       d_tw2, d_itw2, c.gen);
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
