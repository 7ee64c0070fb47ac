// Device-side Fiat-Shamir step of the FRI commit cascade:
//   MerkleChannel.mixRoot (src/core/vcs_lifted/blake2_merkle.zig:237-249):
//       digest <- Blake2s(digest || root), n_draws <- 0
//   Blake2sChannel.drawSecureFelt (src/core/channel/blake2s.zig:41-134,260-279):
//       words = Blake2s(digest || LE32(n_draws) || 0x00), n_draws += 1, redrawn
//       until all 8 words are < 2p; alpha = the first 4 words mod p.
// Run by one thread; the result lands in the DevChannel in HBM.
#pragma once
#include "blake2s_quad.hpp"
#include "runtime.hpp"

namespace sr {

__device__ __forceinline__ void chan_mix_root_draw_thread(DevChannel* __restrict__ ch, const uint32_t root[8],
                                                          uint32_t* __restrict__ root_out) {
    if (root_out) {
        *(uint4*)root_out = make_uint4(root[0], root[1], root[2], root[3]);
        *(uint4*)(root_out + 4) = make_uint4(root[4], root[5], root[6], root[7]);
    }
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = ch->digest[i];
        m[8 + i] = root[i];
    }
    uint32_t h[8];
    b2_init(h);
    b2_compress(h, m, 64, true);
    uint32_t n_draws = 0;
    for (int attempt = 0; attempt < 64; attempt++) {
        uint32_t w[16] = {h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], n_draws, 0, 0, 0, 0, 0, 0, 0};
        uint32_t o[8];
        b2_init(o);
        b2_compress(o, w, 37, true);
        n_draws++;
        bool ok = true;
#pragma unroll
        for (int i = 0; i < 8; i++) ok = ok && (o[i] < 2u * P);
        if (ok) {
            qm31 a = {{mred64(o[0]), mred64(o[1])}, {mred64(o[2]), mred64(o[3])}};
            qm31 a2 = qsqr(a);
#pragma unroll
            for (int i = 0; i < 8; i++) ch->digest[i] = h[i];
            ch->n_draws = n_draws;
            qstore(a, ch->alpha);
            qstore(a2, ch->alpha2);
            return;
        }
    }
    ch->status = 1;  // 64 consecutive rejections (probability ~2^-1600): reported to the host
}

// The same step run by the four lanes of a quad (all active), the root given
// as quad-distributed words (rlo, rhi) = (root[q], root[4+q]).
__device__ __forceinline__ void chan_mix_root_draw_quad(DevChannel* __restrict__ ch, uint32_t rlo, uint32_t rhi,
                                                        uint32_t* __restrict__ root_out) {
    const uint32_t q = __lane_id() & 3;
    const QuadLane L(q);
    if (root_out) {
        root_out[q] = rlo;
        root_out[4 + q] = rhi;
    }
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) m[i] = ch->digest[i];
    quad_gather8(rlo, rhi, m + 8);
    const uint32_t ilo = L.sel(SR_B2_IV0 ^ 0x01010020u, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    const uint32_t ihi = L.sel(SR_B2_IV4, SR_B2_IV5, SR_B2_IV6, SR_B2_IV7);
    uint32_t hlo = ilo, hhi = ihi;
    b2_compress_quad(L, hlo, hhi, m, 64, true);
    uint32_t w[16] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    quad_gather8(hlo, hhi, w);
    for (uint32_t n_draws = 0; n_draws < 64; n_draws++) {
        w[8] = n_draws;
        uint32_t olo = ilo, ohi = ihi;
        b2_compress_quad(L, olo, ohi, w, 37, true);
        uint32_t bad = (olo >= 2u * P || ohi >= 2u * P) ? 1u : 0u;
        bad |= quad_perm<QP_ROT1>(bad);
        bad |= quad_perm<QP_ROT2>(bad);  // OR over the quad
        if (!bad) {
            qm31 a = {{mred64(quad_bcast<0>(olo)), mred64(quad_bcast<1>(olo))},
                      {mred64(quad_bcast<2>(olo)), mred64(quad_bcast<3>(olo))}};
            ch->digest[q] = hlo;
            ch->digest[4 + q] = hhi;
            if (q == 0) {
                ch->n_draws = n_draws + 1;
                qstore(a, ch->alpha);
                qstore(qsqr(a), ch->alpha2);
            }
            return;
        }
    }
    if (q == 0) ch->status = 1;
}

// General channel steps on a quad whose four lanes all hold the digest dg[8]
// (Blake2sChannel.mixFelts / drawSecureFelt, src/core/channel/blake2s.zig:52-134).
// digest <- Blake2s(digest || pay[0..n_words)) for n_words <= 16; n_draws restarts at 0.
__device__ __forceinline__ void qchan_mix(const QuadLane& L, uint32_t dg[8], const uint32_t pay[16], uint32_t n_words) {
    uint32_t hlo = L.sel(SR_B2_IV0 ^ 0x01010020u, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    uint32_t hhi = L.sel(SR_B2_IV4, SR_B2_IV5, SR_B2_IV6, SR_B2_IV7);
    const uint32_t len = 32 + 4 * n_words;
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        m[i] = dg[i];
        m[8 + i] = (uint32_t)i < n_words ? pay[i] : 0u;
    }
    if (len <= 64) {
        b2_compress_quad(L, hlo, hhi, m, len, true);
    } else {
        b2_compress_quad(L, hlo, hhi, m, 64, false);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            m[i] = (uint32_t)(8 + i) < n_words ? pay[8 + i] : 0u;
            m[8 + i] = 0;
        }
        b2_compress_quad(L, hlo, hhi, m, len, true);
    }
    quad_gather8(hlo, hhi, dg);
}

// Draws one secure felt: words = Blake2s(digest || LE32(n_draws) || 0x00), n_draws += 1,
// redrawn until all 8 words are < 2p.  At most 64 attempts; `status` is set to 1 after
// 64 rejections (probability ~2^-1600).
__device__ __forceinline__ qm31 qchan_draw(const QuadLane& L, const uint32_t dg[8], uint32_t& n_draws,
                                           uint32_t& status) {
    const uint32_t ilo = L.sel(SR_B2_IV0 ^ 0x01010020u, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    const uint32_t ihi = L.sel(SR_B2_IV4, SR_B2_IV5, SR_B2_IV6, SR_B2_IV7);
    uint32_t w[16] = {dg[0], dg[1], dg[2], dg[3], dg[4], dg[5], dg[6], dg[7], 0, 0, 0, 0, 0, 0, 0, 0};
    for (int attempt = 0; attempt < 64; attempt++) {
        w[8] = n_draws;
        uint32_t olo = ilo, ohi = ihi;
        b2_compress_quad(L, olo, ohi, w, 37, true);
        n_draws++;
        uint32_t bad = (olo >= 2u * P || ohi >= 2u * P) ? 1u : 0u;
        bad |= quad_perm<QP_ROT1>(bad);
        bad |= quad_perm<QP_ROT2>(bad);
        if (!bad)
            return {{mred64(quad_bcast<0>(olo)), mred64(quad_bcast<1>(olo))},
                    {mred64(quad_bcast<2>(olo)), mred64(quad_bcast<3>(olo))}};
    }
    status = 1;
    return qzero();
}

// QM31 value broadcast from quad lane K to the whole quad.
template <int K>
__device__ __forceinline__ qm31 quad_bcast_q(qm31 v) {
    return {{quad_bcast<K>(v.c0.a), quad_bcast<K>(v.c0.b)}, {quad_bcast<K>(v.c1.a), quad_bcast<K>(v.c1.b)}};
}

}  // namespace sr
