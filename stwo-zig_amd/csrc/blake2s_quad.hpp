// One Blake2s compression spread over the four lanes of a quad (device only).
//
// Lane q of the quad (q = lane & 3) holds column q of the 4x4 working state:
// a = v[q], b = v[4+q], c = v[8+q], d = v[12+q], and the chaining words h[q],
// h[4+q].  The column step is then one G per lane; for the diagonal step rows
// b, c, d are rotated by 1, 2, 3 lanes with DPP quad permutes (lane q reads
// lane q+k), which lines up (v[q], v[4+(q+1)%4], v[8+(q+2)%4], v[12+(q+3)%4]),
// and rotated back afterwards.  Each lane picks its two message words per step
// with lane-mask selects.  A serial chain of compressions (the top levels of a
// Merkle tree, the channel) then advances ~2.5x faster than on one lane.
// All four lanes of a quad must be active together.
#pragma once
#include "blake2s.hpp"

namespace sr {

// quad_perm controls: lane q of each quad reads lane (q + k) & 3.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, 0xF, 0xF, false);
}
constexpr int QP_ROT1 = 0x39, QP_ROT2 = 0x4E, QP_ROT3 = 0x93;
// quad_perm [k,k,k,k]: every lane reads lane k.
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t x) {
    return quad_perm<K * 0x55>(x);
}

struct QuadLane {
    bool q0, q1, q2;
    __device__ __forceinline__ explicit QuadLane(uint32_t q) : q0(q == 0), q1(q == 1), q2(q == 2) {}
    __device__ __forceinline__ uint32_t sel(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) const {
        return q0 ? w0 : (q1 ? w1 : (q2 ? w2 : w3));
    }
};

#define SR_B2Q_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
    {                                                                                     \
        uint32_t x = L.sel(m[s0], m[s2], m[s4], m[s6]), y = L.sel(m[s1], m[s3], m[s5], m[s7]); \
        SR_B2_G(a, b, c, d, x, y)                                                         \
        b = quad_perm<QP_ROT1>(b);                                                        \
        c = quad_perm<QP_ROT2>(c);                                                        \
        d = quad_perm<QP_ROT3>(d);                                                        \
        x = L.sel(m[s8], m[s10], m[s12], m[s14]);                                         \
        y = L.sel(m[s9], m[s11], m[s13], m[s15]);                                         \
        SR_B2_G(a, b, c, d, x, y)                                                         \
        b = quad_perm<QP_ROT3>(b);                                                        \
        c = quad_perm<QP_ROT2>(c);                                                        \
        d = quad_perm<QP_ROT1>(d);                                                        \
    }

// (hlo, hhi) = chaining words (q, 4+q) in and out; m = the 16 message words
// (the same in all four lanes); t = byte counter after this block.
__device__ __forceinline__ void b2_compress_quad(const QuadLane& L, uint32_t& hlo, uint32_t& hhi,
                                                 const uint32_t m[16], uint32_t t, bool last) {
    uint32_t a = hlo, b = hhi;
    uint32_t c = L.sel(SR_B2_IV0, SR_B2_IV1, SR_B2_IV2, SR_B2_IV3);
    uint32_t d = L.sel(SR_B2_IV4 ^ t, SR_B2_IV5, last ? ~SR_B2_IV6 : SR_B2_IV6, SR_B2_IV7);
    SR_B2Q_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
    SR_B2Q_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
    SR_B2Q_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
    SR_B2Q_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
    SR_B2Q_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
    SR_B2Q_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
    SR_B2Q_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
    SR_B2Q_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
    SR_B2Q_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
    SR_B2Q_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
    hlo ^= a ^ c;
    hhi ^= b ^ d;
}

// The 8 words of a quad-distributed hash, in every lane of the quad.
__device__ __forceinline__ void quad_gather8(uint32_t hlo, uint32_t hhi, uint32_t out[8]) {
    out[0] = quad_bcast<0>(hlo);
    out[1] = quad_bcast<1>(hlo);
    out[2] = quad_bcast<2>(hlo);
    out[3] = quad_bcast<3>(hlo);
    out[4] = quad_bcast<0>(hhi);
    out[5] = quad_bcast<1>(hhi);
    out[6] = quad_bcast<2>(hhi);
    out[7] = quad_bcast<3>(hhi);
}

}  // namespace sr
