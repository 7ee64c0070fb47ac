// Transcript steps between the commitments of a proof, on the device channel.
//
// The prover's Fiat-Shamir order (prove.zig:209-389; scheme.zig:571-800) between the
// main-trace commitment and the FRI commitment is a fixed sequence of mixes and draws
// whose results only parameterise the next kernels.  Running it on the device (one quad
// of lanes hashes, chan.hpp) removes the host round trips: the host enqueues the whole
// sequence and reads the roots and sampled values back once, with the FRI state.
//   tr_commit_draw : mixRoot(root) [+ mixU32s(statement)] ; alpha = drawSecureFelt ;
//                    powers[k] = alpha^k                      (prove.zig:237-318)
//   tr_oods        : mixRoot(root) ; t = drawSecureFelt ; OODS point
//                    ((1-t^2)/(1+t^2), 2t/(1+t^2)) (circle.zig:159-167) and the point
//                    evaluation factors [y, x, 2x^2-1, ...] (point_evaluation.zig:14-38)
//   tr_quotient    : mixFelts(sampled values) (sampled_value_transcript.mixIntoChannel) ;
//                    q = drawSecureFelt ; the one-batch quotient coefficients of
//                    samples.zig:245-312 (all samples at the OODS point)
#include "chan.hpp"
#include "prover.hpp"

namespace sr {

namespace {

// digest <- Blake2s(digest || words[0..n)) for any n: the streaming hash over 16-word blocks.
__device__ void qchan_mix_long(const QuadLane& L, uint32_t dg[8], const uint32_t* __restrict__ words, uint32_t n) {
    uint32_t hlo = L.sel(SR_B2_IV0 ^ 0x01010020u, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    uint32_t hhi = L.sel(SR_B2_IV4, SR_B2_IV5, SR_B2_IV6, SR_B2_IV7);
    const uint32_t total = 8 + n;  // words
    const uint32_t nblk = (total + 15) / 16;
    // Block b >= 1 holds words [16 b - 8, 16 b + 8) of `words`.  The next block's words are
    // loaded while a block is compressed (the chain waits on no load), and every block before
    // the last is full, so only the last one checks the bounds.
    auto load = [&](uint32_t b, uint32_t(&dst)[16]) {
        const uint32_t o = 16 * b - 8;
        if (b + 1 < nblk) {
#pragma unroll
            for (int i = 0; i < 16; i++) dst[i] = words[o + i];
        } else {
#pragma unroll
            for (int i = 0; i < 16; i++) dst[i] = o + i < n ? words[o + i] : 0u;
        }
    };
    uint32_t m[16], nx[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = dg[i];
        m[8 + i] = (uint32_t)i < n ? words[i] : 0u;
    }
    if (nblk > 1) load(1, nx);
    b2_compress_quad(L, hlo, hhi, m, nblk == 1 ? 4 * total : 64, nblk == 1);
    for (uint32_t b = 1; b < nblk; b++) {
#pragma unroll
        for (int i = 0; i < 16; i++) m[i] = nx[i];
        if (b + 1 < nblk) load(b + 1, nx);
        const bool last = b + 1 == nblk;
        b2_compress_quad(L, hlo, hhi, m, last ? 4 * total : 64 * (b + 1), last);
    }
    quad_gather8(hlo, hhi, dg);
}

__device__ __forceinline__ qm31 tr_pow(qm31 x, uint32_t e) {
    qm31 r = qone();
    while (e) {
        if (e & 1) r = qmul(r, x);
        x = qsqr(x);
        e >>= 1;
    }
    return r;
}

// Proof blockIdx.y of a batch: its buffers bstride bytes after proof 0's.
template <class T>
__device__ __forceinline__ T* tr_boff(T* p, uint64_t bstride) {
    return p ? (T*)((uint8_t*)p + (uint64_t)blockIdx.y * bstride) : p;
}
template <class T>
__device__ __forceinline__ const T* tr_boff(const T* p, uint64_t bstride) {
    return p ? (const T*)((const uint8_t*)p + (uint64_t)blockIdx.y * bstride) : p;
}

__device__ __forceinline__ void tr_load_dg(const DevChannel* ch, uint32_t dg[8]) {
#pragma unroll
    for (int k = 0; k < 8; k++) dg[k] = ch->digest[k];
}
__device__ __forceinline__ void tr_store(DevChannel* ch, const uint32_t dg[8], uint32_t nd, uint32_t st) {
#pragma unroll
    for (int k = 0; k < 8; k++) ch->digest[k] = dg[k];
    ch->n_draws = nd;
    if (st) ch->status = 1;
}

__global__ __launch_bounds__(256) void tr_commit_draw_kernel(DevChannel* __restrict__ ch, const uint32_t* __restrict__ root,
                                                             uint32_t w0, uint32_t w1, uint32_t n_words,
                                                             uint4* __restrict__ powers, uint32_t n_pow,
                                                             uint64_t bstride) {
    __shared__ uint32_t a_sh[4];
    const uint32_t t = threadIdx.x;
    ch = tr_boff(ch, bstride);
    root = tr_boff(root, bstride);
    powers = tr_boff(powers, bstride);
    if (t < 4) {
        const QuadLane Lq(t);
        uint32_t dg[8], pay[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        tr_load_dg(ch, dg);
#pragma unroll
        for (int k = 0; k < 8; k++) pay[k] = root[k];
        qchan_mix(Lq, dg, pay, 8);
        if (n_words) {
            pay[0] = w0;
            pay[1] = w1;
            qchan_mix(Lq, dg, pay, n_words);
        }
        uint32_t nd = 0, st = 0;
        const qm31 a = qchan_draw(Lq, dg, nd, st);
        if (t == 0) {
            tr_store(ch, dg, nd, st);
            qstore(a, a_sh);
        }
    }
    __syncthreads();
    const qm31 alpha = qload(a_sh);
    for (uint32_t k = t; k < n_pow; k += blockDim.x) {
        const qm31 v = tr_pow(alpha, k);
        powers[k] = make_uint4(v.c0.a, v.c0.b, v.c1.a, v.c1.b);
    }
}

// root == nullptr: the point is drawn without a mix first (the draw counter continues).
__global__ void tr_oods_kernel(DevChannel* __restrict__ ch, const uint32_t* __restrict__ root, uint32_t n_factors,
                               uint4* __restrict__ oods, uint4* __restrict__ factors, uint64_t bstride) {
    const uint32_t t = threadIdx.x;
    if (t >= 4) return;
    ch = tr_boff(ch, bstride);
    root = tr_boff(root, bstride);
    oods = tr_boff(oods, bstride);
    factors = tr_boff(factors, bstride);
    const QuadLane Lq(t);
    uint32_t dg[8], pay[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    tr_load_dg(ch, dg);
    uint32_t nd = ch->n_draws, st = 0;
    if (root) {
#pragma unroll
        for (int k = 0; k < 8; k++) pay[k] = root[k];
        qchan_mix(Lq, dg, pay, 8);
        nd = 0;
    }
    const qm31 tt = qchan_draw(Lq, dg, nd, st);
    if (t != 0) return;
    tr_store(ch, dg, nd, st);
    const qm31 t2 = qsqr(tt);
    const qm31 inv1 = qinv(qadd(t2, qone()));
    const qm31 x = qmul(qsub(qone(), t2), inv1), y = qmul(qadd(tt, tt), inv1);
    oods[0] = make_uint4(x.c0.a, x.c0.b, x.c1.a, x.c1.b);
    oods[1] = make_uint4(y.c0.a, y.c0.b, y.c1.a, y.c1.b);
    qm31 f = y;
    for (uint32_t k = 0; k < n_factors; k++) {
        if (k == 1) f = x;
        if (k >= 2) {
            const qm31 s = qsqr(f);
            f = qsub(qadd(s, s), qone());
        }
        factors[k] = make_uint4(f.c0.a, f.c0.b, f.c1.a, f.c1.b);
    }
}

// digest <- Blake2s(digest || words[0..n_words)) when `mix` (mixRoot, mixFelts and mixU32s are
// all this on little-endian words), then n_draws secure felts into out[0..n_draws).
__global__ void tr_mix_draws_kernel(DevChannel* __restrict__ ch, const uint32_t* __restrict__ words,
                                    uint32_t n_words, uint32_t n_draws, uint4* __restrict__ out, uint32_t mix) {
    const uint32_t t = threadIdx.x;
    if (t >= 4) return;
    const QuadLane Lq(t);
    uint32_t dg[8];
    tr_load_dg(ch, dg);
    uint32_t nd = ch->n_draws, st = 0;
    if (mix) {
        qchan_mix_long(Lq, dg, words, n_words);
        nd = 0;
    }
    for (uint32_t k = 0; k < n_draws; k++) {
        const qm31 a = qchan_draw(Lq, dg, nd, st);
        if (t == 0) out[k] = make_uint4(a.c0.a, a.c0.b, a.c1.a, a.c1.b);
    }
    if (t == 0) tr_store(ch, dg, nd, st);
}

// n_blocks draws of 8 raw words (Blake2sChannel.drawU32s, src/core/channel/blake2s.zig:90-107):
// block k = Blake2s(digest || LE32(n_draws) || 0x00), n_draws += 1 per block, no rejection.
__global__ void tr_draw_words_kernel(DevChannel* __restrict__ ch, uint32_t n_blocks, uint32_t* __restrict__ out) {
    const uint32_t t = threadIdx.x;
    if (t >= 4) return;
    const QuadLane Lq(t);
    const uint32_t ilo = Lq.sel(SR_B2_IV0 ^ 0x01010020u, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    const uint32_t ihi = Lq.sel(SR_B2_IV4, SR_B2_IV5, SR_B2_IV6, SR_B2_IV7);
    uint32_t w[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    tr_load_dg(ch, w);
    uint32_t nd = ch->n_draws;
    for (uint32_t k = 0; k < n_blocks; k++) {
        w[8] = nd++;
        uint32_t olo = ilo, ohi = ihi;
        b2_compress_quad(Lq, olo, ohi, w, 37, true);
        out[8 * k + t] = olo;
        out[8 * k + 4 + t] = ohi;
    }
    if (t == 0) ch->n_draws = nd;
}

// One workgroup: mix the sampled values, draw the quotient coefficient q and fill the
// coefficients of the ns entries (column order) and the single batch at the OODS point:
// entry i: q^i * (conj(py) - py); sum_a = sum q^i (conj(v_i) - v_i);
// sum_b = sum q^i (v_i (conj(py) - py) - (conj(v_i) - v_i) py).
constexpr uint32_t TRQ_THREADS = 256;
constexpr uint32_t TRQ_MAX_NS = 2048;  // sampled values staged in LDS (32 KiB)
__global__ __launch_bounds__(TRQ_THREADS) void tr_quotient_kernel(DevChannel* __restrict__ ch,
                                                                  const uint32_t* __restrict__ sv, uint32_t ns,
                                                                  const uint4* __restrict__ oods,
                                                                  QEntryDev* __restrict__ ents,
                                                                  QuotientBatchDev* __restrict__ bat,
                                                                  uint64_t bstride) {
    __shared__ uint32_t q_sh[4];
    __shared__ qm31 ra[TRQ_THREADS], rb[TRQ_THREADS];
    __shared__ uint32_t svs[4 * TRQ_MAX_NS];
    const uint32_t t = threadIdx.x;
    ch = tr_boff(ch, bstride);
    sv = tr_boff(sv, bstride);
    oods = tr_boff(oods, bstride);
    ents = tr_boff(ents, bstride);
    bat = tr_boff(bat, bstride);
    // the hash reads its message from LDS: one load round trip instead of one per block
    for (uint32_t k = t; k < 4 * ns; k += TRQ_THREADS) svs[k] = sv[k];
    __syncthreads();
    if (t < 4) {
        const QuadLane Lq(t);
        uint32_t dg[8];
        tr_load_dg(ch, dg);
        qchan_mix_long(Lq, dg, svs, 4 * ns);
        uint32_t nd = 0, st = 0;
        const qm31 q = qchan_draw(Lq, dg, nd, st);
        if (t == 0) {
            tr_store(ch, dg, nd, st);
            qstore(q, q_sh);
        }
    }
    __syncthreads();
    const qm31 q = qload(q_sh);
    const uint4 ox = oods[0], oy = oods[1];
    const qm31 px = {{ox.x, ox.y}, {ox.z, ox.w}}, py = {{oy.x, oy.y}, {oy.z, oy.w}};
    const qm31 cc = qsub(qconj(py), py);
    qm31 sa = qzero(), sb = qzero();
    for (uint32_t i = t; i < ns; i += TRQ_THREADS) {
        const qm31 rp = tr_pow(q, i);
        const qm31 v = qload(svs + 4 * i);
        const qm31 a = qsub(qconj(v), v);
        const qm31 b = qsub(qmul(v, cc), qmul(a, py));
        qstore(qmul(rp, cc), ents[i].coeff);
        sa = qadd(sa, qmul(rp, a));
        sb = qadd(sb, qmul(rp, b));
    }
    ra[t] = sa;
    rb[t] = sb;
    __syncthreads();
    for (uint32_t s = TRQ_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) {
            ra[t] = qadd(ra[t], ra[t + s]);
            rb[t] = qadd(rb[t], rb[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        const cm31 prx = px.c0, pry = py.c0, pix = px.c1, piy = py.c1;
        QuotientBatchDev bd;
        bd.det = csub(cmul(prx, piy), cmul(pry, pix));
        bd.pix = pix;
        bd.piy = piy;
        bd.sum_a = ra[0];
        bd.sum_b = rb[0];
        bd.first = 0;
        bd.count = ns;
        *bat = bd;
    }
}

}  // namespace

void launch_tr_commit_draw(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, const uint32_t* words, uint32_t n_words,
                           uint32_t* d_powers, uint32_t n_pow, uint32_t batch, uint64_t bstride) {
    SR_CHECK(n_words <= 2, "transcript: at most 2 statement words");
    SR_CHECK(batch >= 1 && batch <= 65535, "transcript: batch size");
    KTimer kt(c, "transcript", 0.0);
    hipLaunchKernelGGL(tr_commit_draw_kernel, dim3(1, batch), dim3(256), 0, c.stream, d_ch, (const uint32_t*)d_root,
                       n_words > 0 ? words[0] : 0u, n_words > 1 ? words[1] : 0u, n_words, (uint4*)d_powers, n_pow,
                       bstride);
    SR_HIP(hipGetLastError());
}

void launch_tr_mix_draws(Ctx& c, DevChannel* d_ch, const uint32_t* d_words, uint32_t n_words, uint32_t n_draws,
                         uint32_t* d_out, bool mix_empty) {
    SR_CHECK(n_words == 0 || d_words, "transcript: words");
    SR_CHECK(n_draws == 0 || d_out, "transcript: draw output");
    KTimer kt(c, "transcript", 4.0 * n_words);
    hipLaunchKernelGGL(tr_mix_draws_kernel, dim3(1), dim3(64), 0, c.stream, d_ch, d_words, n_words, n_draws,
                       (uint4*)d_out, (n_words || mix_empty) ? 1u : 0u);
    SR_HIP(hipGetLastError());
}

void launch_tr_draw_words(Ctx& c, DevChannel* d_ch, uint32_t n_blocks, uint32_t* d_out) {
    SR_CHECK(n_blocks == 0 || d_out, "transcript: draw output");
    KTimer kt(c, "transcript", 32.0 * n_blocks);
    hipLaunchKernelGGL(tr_draw_words_kernel, dim3(1), dim3(64), 0, c.stream, d_ch, n_blocks, d_out);
    SR_HIP(hipGetLastError());
}

void launch_tr_oods(Ctx& c, DevChannel* d_ch, const uint8_t* d_root, uint32_t n_factors, uint32_t* d_oods,
                    uint32_t* d_factors, uint32_t batch, uint64_t bstride) {
    SR_CHECK(n_factors <= 30, "transcript: too many evaluation factors");
    SR_CHECK(batch >= 1 && batch <= 65535, "transcript: batch size");
    KTimer kt(c, "transcript", 0.0);
    hipLaunchKernelGGL(tr_oods_kernel, dim3(1, batch), dim3(64), 0, c.stream, d_ch, (const uint32_t*)d_root, n_factors,
                       (uint4*)d_oods, (uint4*)d_factors, bstride);
    SR_HIP(hipGetLastError());
}

void launch_tr_quotient(Ctx& c, DevChannel* d_ch, const uint32_t* d_sv, uint32_t ns, const uint32_t* d_oods,
                        QEntryDev* d_ents, QuotientBatchDev* d_bat, uint32_t batch, uint64_t bstride) {
    SR_CHECK(ns >= 1 && ns <= TRQ_MAX_NS, "transcript: sampled value count");
    SR_CHECK(batch >= 1 && batch <= 65535, "transcript: batch size");
    KTimer kt(c, "transcript", batch * 16.0 * ns);
    hipLaunchKernelGGL(tr_quotient_kernel, dim3(1, batch), dim3(TRQ_THREADS), 0, c.stream, d_ch, d_sv, ns,
                       (const uint4*)d_oods, d_ents, d_bat, bstride);
    SR_HIP(hipGetLastError());
}

}  // namespace sr
