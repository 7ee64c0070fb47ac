// Blake2s-256 (RFC 7693, unkeyed, 32-byte digest) for host and device.
//
// Semantics follow src/core/crypto/blake2s_backend.zig:93-160,233-300: the
// counter is the number of message bytes so far, the last block (possibly a
// full one) carries the final flag, and a fixed 64-byte domain prefix can be
// absorbed once into a "seed" state (leaf and node prefixes of
// src/core/vcs_lifted/blake2_merkle.zig:9-11,258-262).
#pragma once
#include <string.h>

#include "fields.hpp"

namespace sr {

struct B2State {
    uint32_t h[8];
};

#define SR_B2_IV0 0x6A09E667u
#define SR_B2_IV1 0xBB67AE85u
#define SR_B2_IV2 0x3C6EF372u
#define SR_B2_IV3 0xA54FF53Au
#define SR_B2_IV4 0x510E527Fu
#define SR_B2_IV5 0x9B05688Cu
#define SR_B2_IV6 0x1F83D9ABu
#define SR_B2_IV7 0x5BE0CD19u

SR_HD uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
// Byte rotations as byte permutes: on gfx950 v_perm_b32 issues at twice the
// rate of v_alignbit_b32 (measured: +7% compressions/s).
SR_HD uint32_t rotr16(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_perm(x, x, 0x01000302u);
#else
    return rotr32(x, 16);
#endif
}
SR_HD uint32_t rotr8(uint32_t x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_perm(x, x, 0x00030201u);
#else
    return rotr32(x, 8);
#endif
}

#define SR_B2_G(a, b, c, d, x, y) \
    a = a + b + (x);              \
    d = rotr16(d ^ a);            \
    c = c + d;                    \
    b = rotr32(b ^ c, 12);        \
    a = a + b + (y);              \
    d = rotr8(d ^ a);             \
    c = c + d;                    \
    b = rotr32(b ^ c, 7);

#define SR_B2_ROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15) \
    SR_B2_G(v0, v4, v8, v12, m[s0], m[s1])                                                \
    SR_B2_G(v1, v5, v9, v13, m[s2], m[s3])                                                \
    SR_B2_G(v2, v6, v10, v14, m[s4], m[s5])                                               \
    SR_B2_G(v3, v7, v11, v15, m[s6], m[s7])                                               \
    SR_B2_G(v0, v5, v10, v15, m[s8], m[s9])                                               \
    SR_B2_G(v1, v6, v11, v12, m[s10], m[s11])                                             \
    SR_B2_G(v2, v7, v8, v13, m[s12], m[s13])                                              \
    SR_B2_G(v3, v4, v9, v14, m[s14], m[s15])

// One compression; t = byte counter after this block (< 2^32 on every path here).
SR_HD void b2_compress(uint32_t h[8], const uint32_t m[16], uint32_t t, bool last) {
    uint32_t v0 = h[0], v1 = h[1], v2 = h[2], v3 = h[3], v4 = h[4], v5 = h[5], v6 = h[6], v7 = h[7];
    uint32_t v8 = SR_B2_IV0, v9 = SR_B2_IV1, v10 = SR_B2_IV2, v11 = SR_B2_IV3;
    uint32_t v12 = SR_B2_IV4 ^ t, v13 = SR_B2_IV5, v14 = last ? ~SR_B2_IV6 : SR_B2_IV6, v15 = SR_B2_IV7;
    SR_B2_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)
    SR_B2_ROUND(14, 10, 4, 8, 9, 15, 13, 6, 1, 12, 0, 2, 11, 7, 5, 3)
    SR_B2_ROUND(11, 8, 12, 0, 5, 2, 15, 13, 10, 14, 3, 6, 7, 1, 9, 4)
    SR_B2_ROUND(7, 9, 3, 1, 13, 12, 11, 14, 2, 6, 5, 10, 4, 0, 15, 8)
    SR_B2_ROUND(9, 0, 5, 7, 2, 4, 10, 15, 14, 1, 11, 12, 6, 8, 3, 13)
    SR_B2_ROUND(2, 12, 6, 10, 0, 11, 8, 3, 4, 13, 7, 5, 15, 14, 1, 9)
    SR_B2_ROUND(12, 5, 1, 15, 14, 13, 4, 10, 0, 7, 6, 3, 9, 2, 8, 11)
    SR_B2_ROUND(13, 11, 7, 14, 12, 1, 3, 9, 5, 0, 15, 4, 8, 6, 2, 10)
    SR_B2_ROUND(6, 15, 14, 9, 11, 3, 0, 8, 12, 2, 13, 7, 1, 4, 10, 5)
    SR_B2_ROUND(10, 2, 8, 4, 7, 6, 1, 5, 15, 11, 9, 14, 3, 12, 13, 0)
    h[0] ^= v0 ^ v8;
    h[1] ^= v1 ^ v9;
    h[2] ^= v2 ^ v10;
    h[3] ^= v3 ^ v11;
    h[4] ^= v4 ^ v12;
    h[5] ^= v5 ^ v13;
    h[6] ^= v6 ^ v14;
    h[7] ^= v7 ^ v15;
}

SR_HD void b2_init(uint32_t h[8]) {
    h[0] = SR_B2_IV0 ^ 0x01010020u;  // digest 32 bytes, no key, fanout 1, depth 1
    h[1] = SR_B2_IV1;
    h[2] = SR_B2_IV2;
    h[3] = SR_B2_IV3;
    h[4] = SR_B2_IV4;
    h[5] = SR_B2_IV5;
    h[6] = SR_B2_IV6;
    h[7] = SR_B2_IV7;
}

// Host-side streaming hasher (channel, prefixes).
struct B2Stream {
    uint32_t h[8];
    uint8_t buf[64];
    uint32_t buflen = 0;
    uint32_t t = 0;
    B2Stream() { b2_init(h); }
    void update(const void* data, size_t n) {
        const uint8_t* d = (const uint8_t*)data;
        while (n > 0) {
            if (buflen == 64) {
                t += 64;
                uint32_t m[16];
                memcpy(m, buf, 64);
                b2_compress(h, m, t, false);
                buflen = 0;
            }
            size_t take = 64 - buflen;
            if (take > n) take = n;
            memcpy(buf + buflen, d, take);
            buflen += (uint32_t)take;
            d += take;
            n -= take;
        }
    }
    void finalize(uint8_t out[32]) {
        t += buflen;
        memset(buf + buflen, 0, 64 - buflen);
        uint32_t m[16];
        memcpy(m, buf, 64);
        b2_compress(h, m, t, true);
        memcpy(out, h, 32);
    }
};

inline void b2_hash(const void* data, size_t n, uint8_t out[32]) {
    B2Stream s;
    s.update(data, n);
    s.finalize(out);
}

// State after absorbing a 64-byte tag prefix "tag"+0^(64-len) as a non-final block.
inline B2State b2_prefix_seed(const char* tag) {
    uint32_t m[16] = {0};
    memcpy(m, tag, strlen(tag));
    B2State s;
    b2_init(s.h);
    b2_compress(s.h, m, 64, false);
    return s;
}

}  // namespace sr
