// Row-parallel proving kernels on gfx950: quotient accumulation (+ fused FRI
// first-layer leaf hash), wide-Fibonacci composition and trace generation,
// and sampled-value point evaluation.
//
// All are HBM-streaming kernels: each reads its input columns once with
// coalesced loads (consecutive lanes = consecutive rows of one column) and
// accumulates M31 products in 64-bit registers (one Mersenne fold per product,
// one canonicalisation per sum), which is exact and order independent.
#include <vector>

#include "runtime.hpp"

namespace sr {

// Quotients (core/pcs/quotients/row_evaluation.zig:402-521): for row p with
// domain point (x, y) = domain.at(bitrev(p)):
//   q(p) = sum_b ( sum_{e in b} c_e * v_e(p) - (sum_a_b * y + sum_b_b) ) / den_b(p),
//   den_b = det_b - piy_b * x + pix_b * y  (CM31).
// Entries are flat 32-byte descriptors (column, lift shift, coefficient) read
// with scalar loads; QB column values are loaded at a time so that QB loads
// are in flight per wave, and four raw products are summed before a fold.

__device__ __forceinline__ uint32_t lift_row(uint32_t row, uint32_t shift) {
    return ((row >> shift) << 1) + (row & 1);
}

// Word `i` of an entry's column for the proof at byte offset boff of a batch.
__device__ __forceinline__ uint32_t ent_word(const QEntryDev& en, uint64_t boff, uint32_t i) {
    return ((gcu32*)((const uint8_t*)en.col + boff))[i];
}

template <class T>
__device__ __forceinline__ T* boffp(T* p, uint64_t boff) {
    return (T*)((uint8_t*)p + boff);
}
template <class T>
__device__ __forceinline__ const T* boffp(const T* p, uint64_t boff) {
    return (const T*)((const uint8_t*)p + boff);
}

// The numerator of batch bt at `row`: sum over its entries of coeff * column value, minus the
// batch's linear term.  IDENT: every entry's column has the domain's size (lift(row) = row), so
// each word is the entry's column pointer (uniform) plus the lane's byte offset row * 4.
template <bool IDENT>
__device__ __forceinline__ qm31 q_batch_num(const QEntryDev* __restrict__ entries, const QuotientBatchDev& bt,
                                            uint32_t row, uint32_t y, uint64_t boff) {
    constexpr uint32_t QB = 8;
    const uint32_t voff = row << 2;
    auto word = [&](const QEntryDev& en) -> uint32_t {
        if constexpr (IDENT) return *(gcu32*)((const SR_GLOBAL uint8_t*)en.col + boff + voff);
        else return ent_word(en, boff, lift_row(row, en.shift));
    };
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    uint32_t e = bt.first;
    const uint32_t end = bt.first + bt.count;
    for (; e + QB <= end; e += QB) {
        uint32_t v[QB];
#pragma unroll
        for (uint32_t k = 0; k < QB; k++) v[k] = word(entries[e + k]);
#pragma unroll
        for (uint32_t h = 0; h < QB; h += 4) {
            uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
            for (uint32_t k = h; k < h + 4; k++) {
                const QEntryDev& en = entries[e + k];
                t0 += (uint64_t)en.coeff[0] * v[k];
                t1 += (uint64_t)en.coeff[1] * v[k];
                t2 += (uint64_t)en.coeff[2] * v[k];
                t3 += (uint64_t)en.coeff[3] * v[k];
            }
            s0 += mfold(t0);
            s1 += mfold(t1);
            s2 += mfold(t2);
            s3 += mfold(t3);
        }
    }
    for (; e < end; e++) {
        const QEntryDev& en = entries[e];
        const uint32_t v = word(en);
        s0 += mfold((uint64_t)en.coeff[0] * v);
        s1 += mfold((uint64_t)en.coeff[1] * v);
        s2 += mfold((uint64_t)en.coeff[2] * v);
        s3 += mfold((uint64_t)en.coeff[3] * v);
    }
    const qm31 num = {{mred64(s0), mred64(s1)}, {mred64(s2), mred64(s3)}};
    return qsub(num, qadd(qmulm(bt.sum_a, y), bt.sum_b));
}

// Quotient row of every batch: acc = sum_b num_b / den_b.  Up to Q_INV_BATCHES batches share
// one CM31 inversion (Montgomery's trick over the row's denominators: prefix products, one
// inverse, then each inverse from the running inverse and the prefix before it); more
// batches invert each denominator.  Field arithmetic is exact, so both forms give the same row.
constexpr uint32_t Q_INV_BATCHES = 4;
template <bool IDENT>
__global__ __launch_bounds__(256) void quotients_kernel(const QEntryDev* __restrict__ entries,
                                                        const QuotientBatchDev* __restrict__ batches,
                                                        uint32_t n_batches, uint32_t log_size,
                                                        const uint32_t* __restrict__ domain_xy,
                                                        uint32_t* __restrict__ o0, uint32_t* __restrict__ o1,
                                                        uint32_t* __restrict__ o2, uint32_t* __restrict__ o3,
                                                        uint8_t* __restrict__ leaves, B2State seed, uint64_t bstride) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (row >= n) return;
    // proof blockIdx.y of a batch: every per-proof buffer (and every column an entry names) sits
    // bstride bytes after the previous proof's
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    entries = (const QEntryDev*)((const uint8_t*)entries + boff);
    batches = (const QuotientBatchDev*)((const uint8_t*)batches + boff);
    const uint32_t x = domain_xy[row], y = domain_xy[n + row];
    qm31 acc = qzero();
    if (n_batches <= Q_INV_BATCHES) {
        qm31 num[Q_INV_BATCHES];
        cm31 den[Q_INV_BATCHES], pre[Q_INV_BATCHES];
#pragma unroll
        for (uint32_t bi = 0; bi < Q_INV_BATCHES; bi++) {
            if (bi < n_batches) {
                const QuotientBatchDev& bt = batches[bi];
                num[bi] = q_batch_num<IDENT>(entries, bt, row, y, boff);
                den[bi] = cadd(csub(bt.det, cmulm(bt.piy, x)), cmulm(bt.pix, y));
                pre[bi] = bi == 0 ? den[bi] : cmul(pre[bi - 1], den[bi]);
            } else {
                pre[bi] = pre[bi - 1];  // bi >= 1: n_batches >= 1
            }
        }
        cm31 inv = cinv(pre[Q_INV_BATCHES - 1]);  // the inverse of the product of all denominators
#pragma unroll
        for (int bi = (int)Q_INV_BATCHES - 1; bi >= 0; bi--) {
            if ((uint32_t)bi < n_batches) {
                const cm31 ib = bi == 0 ? inv : cmul(inv, pre[bi - 1]);
                if (bi > 0) inv = cmul(inv, den[bi]);
                acc = qadd(acc, qmulc(num[bi], ib));
            }
        }
    } else {
        for (uint32_t bi = 0; bi < n_batches; bi++) {
            const QuotientBatchDev bt = batches[bi];
            const qm31 num = q_batch_num<IDENT>(entries, bt, row, y, boff);
            const cm31 den = cadd(csub(bt.det, cmulm(bt.piy, x)), cmulm(bt.pix, y));
            acc = qadd(acc, qmulc(num, cinv(den)));
        }
    }
    boffp(o0, boff)[row] = acc.c0.a;
    boffp(o1, boff)[row] = acc.c0.b;
    boffp(o2, boff)[row] = acc.c1.a;
    boffp(o3, boff)[row] = acc.c1.b;
    if (leaves) {
        leaves += boff;
        uint32_t m[16] = {acc.c0.a, acc.c0.b, acc.c1.a, acc.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t h[8];
#pragma unroll
        for (int k = 0; k < 8; k++) h[k] = seed.h[k];
        b2_compress(h, m, 80, true);
        uint4* o = (uint4*)(leaves + (uint64_t)row * 32);
        o[0] = make_uint4(h[0], h[1], h[2], h[3]);
        o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    }
}

// Few rows and many entries (a wide trace on a small domain): one thread per row would loop
// over every entry with a handful of loads in flight.  Blocks instead take (row block, chunk
// of QS_CHUNK entries) and store their canonical partial sums; a finishing kernel adds a
// row's partials and applies the batch's subtraction, division and leaf hash.  One batch.
constexpr uint32_t QS_CHUNK = 64;
__global__ __launch_bounds__(256) void quotients_split_kernel(const QEntryDev* __restrict__ entries,
                                                              const QuotientBatchDev* __restrict__ batches,
                                                              uint32_t log_size, uint32_t* __restrict__ part,
                                                              uint64_t bstride, uint64_t pstride) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (row >= n) return;
    // proof blockIdx.z of a batch (blockIdx.y is the entry chunk); partials pstride apart
    const uint64_t boff = (uint64_t)blockIdx.z * bstride;
    entries = boffp(entries, boff);
    part = boffp(part, (uint64_t)blockIdx.z * pstride);
    const QuotientBatchDev bt = boffp(batches, boff)[0];
    const uint32_t e0 = bt.first + blockIdx.y * QS_CHUNK;
    const uint32_t end = min(e0 + QS_CHUNK, bt.first + bt.count);
    constexpr uint32_t QB = 8;
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    uint32_t e = e0;
    for (; e + QB <= end; e += QB) {
        uint32_t v[QB];
#pragma unroll
        for (uint32_t k = 0; k < QB; k++) {
            const QEntryDev& en = entries[e + k];
            v[k] = ent_word(en, boff, lift_row(row, en.shift));
        }
#pragma unroll
        for (uint32_t h = 0; h < QB; h += 4) {
            uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
            for (uint32_t k = h; k < h + 4; k++) {
                const QEntryDev& en = entries[e + k];
                t0 += (uint64_t)en.coeff[0] * v[k];
                t1 += (uint64_t)en.coeff[1] * v[k];
                t2 += (uint64_t)en.coeff[2] * v[k];
                t3 += (uint64_t)en.coeff[3] * v[k];
            }
            s0 += mfold(t0);
            s1 += mfold(t1);
            s2 += mfold(t2);
            s3 += mfold(t3);
        }
    }
    for (; e < end; e++) {
        const QEntryDev& en = entries[e];
        const uint32_t v = ent_word(en, boff, lift_row(row, en.shift));
        s0 += mfold((uint64_t)en.coeff[0] * v);
        s1 += mfold((uint64_t)en.coeff[1] * v);
        s2 += mfold((uint64_t)en.coeff[2] * v);
        s3 += mfold((uint64_t)en.coeff[3] * v);
    }
    uint32_t* o = part + (uint64_t)blockIdx.y * 4 * n + row;
    o[0] = mred64(s0);
    o[n] = mred64(s1);
    o[2 * (uint64_t)n] = mred64(s2);
    o[3 * (uint64_t)n] = mred64(s3);
}
__global__ __launch_bounds__(256) void quotients_finish_kernel(const QuotientBatchDev* __restrict__ batches,
                                                               uint32_t n_chunks, uint32_t log_size,
                                                               const uint32_t* __restrict__ domain_xy,
                                                               const uint32_t* __restrict__ part,
                                                               uint32_t* __restrict__ o0, uint32_t* __restrict__ o1,
                                                               uint32_t* __restrict__ o2, uint32_t* __restrict__ o3,
                                                               uint8_t* __restrict__ leaves, B2State seed,
                                                               uint64_t bstride, uint64_t pstride) {
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (row >= n) return;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    part = boffp(part, (uint64_t)blockIdx.y * pstride);
    batches = boffp(batches, boff);
    const uint32_t x = domain_xy[row], y = domain_xy[n + row];
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;  // n_chunks canonical words: far below 2^64
    for (uint32_t k = 0; k < n_chunks; k++) {
        const uint32_t* p = part + (uint64_t)k * 4 * n + row;
        s0 += p[0];
        s1 += p[n];
        s2 += p[2 * (uint64_t)n];
        s3 += p[3 * (uint64_t)n];
    }
    const QuotientBatchDev bt = batches[0];
    qm31 num = {{mred64(s0), mred64(s1)}, {mred64(s2), mred64(s3)}};
    num = qsub(num, qadd(qmulm(bt.sum_a, y), bt.sum_b));
    const cm31 den = cadd(csub(bt.det, cmulm(bt.piy, x)), cmulm(bt.pix, y));
    const qm31 acc = qmulc(num, cinv(den));
    boffp(o0, boff)[row] = acc.c0.a;
    boffp(o1, boff)[row] = acc.c0.b;
    boffp(o2, boff)[row] = acc.c1.a;
    boffp(o3, boff)[row] = acc.c1.b;
    if (leaves) {
        leaves += boff;
        uint32_t m[16] = {acc.c0.a, acc.c0.b, acc.c1.a, acc.c1.b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t h[8];
#pragma unroll
        for (int k = 0; k < 8; k++) h[k] = seed.h[k];
        b2_compress(h, m, 80, true);
        uint4* o = (uint4*)(leaves + (uint64_t)row * 32);
        o[0] = make_uint4(h[0], h[1], h[2], h[3]);
        o[1] = make_uint4(h[4], h[5], h[6], h[7]);
    }
}

void launch_quotients(Ctx& c, const uint32_t* const* h_cols, const uint32_t* h_col_logs, const uint32_t* h_entry_cols,
                      const uint32_t* h_entry_coeffs, uint32_t n_entries, const QuotientBatchDev* h_batches,
                      uint32_t n_batches, uint32_t log_size, uint32_t* const* out, uint8_t* d_leaf_out) {
    std::vector<QEntryDev> ents(n_entries);
    double in_bytes = 0;
    bool ident = true;
    for (uint32_t e = 0; e < n_entries; e++) {
        const uint32_t col = h_entry_cols[e];
        SR_CHECK(h_col_logs[col] >= 1 && h_col_logs[col] <= log_size, "quotients: column larger than the domain");
        ents[e].col = h_cols[col];
        ents[e].shift = log_size - h_col_logs[col] + 1;
        ents[e].pad = 0;
        ident = ident && ents[e].shift == 1;
        for (int k = 0; k < 4; k++) ents[e].coeff[k] = h_entry_coeffs[4 * e + k];
        in_bytes += 4.0 * (double)(1ull << h_col_logs[col]);
    }
    for (uint32_t b = 0; b < n_batches; b++)
        SR_CHECK((uint64_t)h_batches[b].first + h_batches[b].count <= n_entries, "quotients: batch range");
    DBuf d_ent = upload(c, ents.data(), ents.size());
    DBuf d_bat = upload(c, h_batches, n_batches);
    launch_quotients_dev(c, d_ent.as<QEntryDev>(), n_entries, d_bat.as<QuotientBatchDev>(), n_batches, in_bytes,
                         log_size, out, d_leaf_out, 1, 0, ident);
}

void launch_quotients_dev(Ctx& c, const QEntryDev* d_ent, uint32_t n_entries, const QuotientBatchDev* d_bat,
                          uint32_t n_batches, double in_bytes, uint32_t log_size, uint32_t* const* out,
                          uint8_t* d_leaf_out, uint32_t batch, uint64_t bstride, bool ident) {
    SR_CHECK(n_entries >= 1 && n_batches >= 1, "quotients: no samples");
    SR_CHECK(batch >= 1 && batch <= 65535, "quotients: batch size");
    const uint32_t* xy = c.domain_points(log_size);
    uint32_t n = 1u << log_size;
    KTimer kt(c, "quotients", batch * (in_bytes + (8.0 + 16.0 + (d_leaf_out ? 32.0 : 0.0)) * n));
    // the split form pays off when one thread per row leaves most of the device idle
    if (n_batches == 1 && n_entries > 2 * QS_CHUNK && log_size <= 16) {
        const uint32_t n_chunks = (n_entries + QS_CHUNK - 1) / QS_CHUNK;  // the batch spans every entry
        const uint64_t pstride = 16ull * n_chunks * n;
        DBuf part(&c, pstride * batch);
        hipLaunchKernelGGL(quotients_split_kernel, dim3((n + 255) / 256, n_chunks, batch), dim3(256), 0, c.stream,
                           d_ent, d_bat, log_size, part.as<uint32_t>(), bstride, pstride);
        hipLaunchKernelGGL(quotients_finish_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, d_bat,
                           n_chunks, log_size, xy, (const uint32_t*)part.as<uint32_t>(), out[0], out[1], out[2],
                           out[3], d_leaf_out, c.leaf_seed, bstride, pstride);
    } else {
        if (ident)
            hipLaunchKernelGGL(quotients_kernel<true>, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, d_ent,
                               d_bat, n_batches, log_size, xy, out[0], out[1], out[2], out[3], d_leaf_out,
                               c.leaf_seed, bstride);
        else
            hipLaunchKernelGGL(quotients_kernel<false>, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, d_ent,
                               d_bat, n_batches, log_size, xy, out[0], out[1], out[2], out[3], d_leaf_out,
                               c.leaf_seed, bstride);
    }
    SR_HIP(hipGetLastError());
}

// Wide-Fibonacci composition (examples/wide_fibonacci/component.zig:146-228):
//   row r: sum_{i} alpha^(N-1-i) * (c_{i+2} - c_i^2 - c_{i+1}^2) * Z^-1[r >> log_n]
// on CanonicCoset(log_n + 1), N = n_cols - 2 constraints.
__global__ __launch_bounds__(256) void widefib_composition_kernel(const uint32_t* const* __restrict__ cols,
                                                                  uint32_t n_cols, uint32_t log_eval,
                                                                  uint32_t log_n, const uint4* __restrict__ powers,
                                                                  uint32_t den0, uint32_t den1,
                                                                  uint32_t* __restrict__ o0, uint32_t* __restrict__ o1,
                                                                  uint32_t* __restrict__ o2,
                                                                  uint32_t* __restrict__ o3, uint64_t bstride) {
    uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= (1u << log_eval)) return;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;  // proof blockIdx.y of a batch
    powers = boffp(powers, boff);
#define WF_COL(i) ((gcu32*)((const uint8_t*)cols[i] + boff))
    uint32_t a = WF_COL(0)[row], b = WF_COL(1)[row];
    // sa, sb: squares of the two previous columns (each column is squared once).
    uint32_t sa = msqr(a), sb = msqr(b);
    uint64_t s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    const uint32_t nc = n_cols - 2;
    // Column values are loaded CB at a time so that CB loads are in flight per
    // wave; four raw products (< 2^62 each) are summed in 64 bits before one
    // Mersenne fold.
    constexpr uint32_t CB = 8;
    uint32_t i = 0;
    for (; i + CB <= nc; i += CB) {
        uint32_t v[CB];
#pragma unroll
        for (uint32_t k = 0; k < CB; k++) v[k] = WF_COL(i + 2 + k)[row];
#pragma unroll
        for (uint32_t h = 0; h < CB; h += 4) {
            uint64_t t0 = 0, t1 = 0, t2 = 0, t3 = 0;
#pragma unroll
            for (uint32_t k = h; k < h + 4; k++) {
                const uint32_t cons = msub(v[k], madd(sa, sb));
                const uint4 pw = powers[nc - 1 - (i + k)];
                t0 += (uint64_t)pw.x * cons;
                t1 += (uint64_t)pw.y * cons;
                t2 += (uint64_t)pw.z * cons;
                t3 += (uint64_t)pw.w * cons;
                sa = sb;
                sb = msqr(v[k]);
            }
            s0 += mfold(t0);
            s1 += mfold(t1);
            s2 += mfold(t2);
            s3 += mfold(t3);
        }
    }
    for (; i < nc; i++) {
        const uint32_t cv = WF_COL(i + 2)[row];
        const uint32_t cons = msub(cv, madd(sa, sb));
        const uint4 pw = powers[nc - 1 - i];
        s0 += mfold((uint64_t)pw.x * cons);
        s1 += mfold((uint64_t)pw.y * cons);
        s2 += mfold((uint64_t)pw.z * cons);
        s3 += mfold((uint64_t)pw.w * cons);
        sa = sb;
        sb = msqr(cv);
    }
#undef WF_COL
    uint32_t d = (row >> log_n) ? den1 : den0;
    boffp(o0, boff)[row] = mmul(mred64(s0), d);
    boffp(o1, boff)[row] = mmul(mred64(s1), d);
    boffp(o2, boff)[row] = mmul(mred64(s2), d);
    boffp(o3, boff)[row] = mmul(mred64(s3), d);
}

void launch_widefib_composition(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_eval,
                                uint32_t log_n, const uint32_t* d_powers, uint32_t den0, uint32_t den1,
                                uint32_t* const* out, uint32_t batch, uint64_t bstride) {
    SR_CHECK(n_cols >= 3, "composition needs at least 3 columns");
    SR_CHECK(batch >= 1 && batch <= 65535, "composition: batch size");
    uint32_t n = 1u << log_eval;
    KTimer kt(c, "composition", batch * (4.0 * n_cols + 16.0) * n);
    hipLaunchKernelGGL(widefib_composition_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, d_cols,
                       n_cols, log_eval, log_n, (const uint4*)d_powers, den0, den1, out[0], out[1], out[2], out[3],
                       bstride);
    SR_HIP(hipGetLastError());
}

// Wide-Fibonacci trace in bit-reversed circle-domain order
// (examples/wide_fibonacci/trace.zig:40-162).  mode 0: (1, logicalRow(s));
// mode 1: seeded uniform M31 pair per storage index.
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__global__ __launch_bounds__(256) void widefib_trace_kernel(uint32_t log_n, uint32_t seq_len, uint32_t mode,
                                                            uint64_t seed, uint32_t* __restrict__ out,
                                                            uint64_t stride, const uint64_t* __restrict__ seeds,
                                                            uint64_t bstride) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t n = 1u << log_n;
    if (s >= n) return;
    // proof blockIdx.y of a batch: its own seed and trace bstride bytes further on
    if (seeds) seed = seeds[blockIdx.y];
    out = boffp(out, (uint64_t)blockIdx.y * bstride);
    uint32_t a, b;
    if (mode == 0) {
        uint32_t ci = bitrev32(s, log_n);
        uint32_t row = ci < n / 2 ? ci * 2 : (n - 1 - ci) * 2 + 1;
        a = 1;
        b = row;
    } else {
        a = (uint32_t)(splitmix64(seed ^ ((2ull * s) * 0xD1B54A32D192ED03ull)) % P);
        b = (uint32_t)(splitmix64(seed ^ ((2ull * s + 1) * 0xD1B54A32D192ED03ull)) % P);
    }
    out[s] = a;
    out[stride + s] = b;
    for (uint32_t col = 2; col < seq_len; col++) {
        uint32_t nx = madd(msqr(a), msqr(b));
        out[(uint64_t)col * stride + s] = nx;
        a = b;
        b = nx;
    }
}
void launch_widefib_trace(Ctx& c, uint32_t log_n, uint32_t seq_len, uint32_t mode, uint64_t seed, uint32_t* d_out,
                          uint64_t col_stride, uint32_t batch, const uint64_t* d_seeds, uint64_t bstride) {
    SR_CHECK(batch >= 1 && batch <= 65535, "widefib trace: batch size");
    uint32_t n = 1u << log_n;
    KTimer kt(c, "widefib_trace", batch * 4.0 * seq_len * n);
    hipLaunchKernelGGL(widefib_trace_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, log_n, seq_len,
                       mode, seed, d_out, col_stride, d_seeds, bstride);
    SR_HIP(hipGetLastError());
}

// Poseidon AIR trace (examples/poseidon/input.zig:121-211): one thread per (row, instance);
// thread t handles row t mod 2^log_n of instance t >> log_n, so consecutive lanes write
// consecutive words of every column.  The instance's columns are its input state, the state
// after each of the 4 + 4 full rounds and state[0] after each of the 14 partial rounds.
constexpr uint32_t POS_COLS_PER_REP = 16 * 9 + 14;
__device__ __forceinline__ uint32_t pos_pow5(uint32_t x) {
    const uint32_t x2 = msqr(x);
    return mmul(msqr(x2), x);
}
// External round (input.zig:140-147,170-201): add constants, M4 per block of four, add the
// column sums of the four blocks, then x^5.
__device__ __forceinline__ void pos_external(uint32_t s[16], uint32_t round) {
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) s[i] = madd(s[i], 1234u + round * 37u + i);
#pragma unroll
    for (uint32_t b = 0; b < 16; b += 4) {
        const uint32_t t0 = madd(s[b], s[b + 1]), t1 = madd(s[b + 2], s[b + 3]);
        const uint32_t t2 = madd(madd(s[b + 1], s[b + 1]), t1);
        const uint32_t t3 = madd(madd(s[b + 3], s[b + 3]), t0);
        const uint32_t t12 = madd(t1, t1), t02 = madd(t0, t0);
        const uint32_t t4 = madd(madd(t12, t12), t3);
        const uint32_t t5 = madd(madd(t02, t02), t2);
        s[b] = madd(t3, t5);
        s[b + 1] = t5;
        s[b + 2] = madd(t2, t4);
        s[b + 3] = t4;
    }
#pragma unroll
    for (uint32_t j = 0; j < 4; j++) {
        const uint32_t sum = madd(madd(s[j], s[j + 4]), madd(s[j + 8], s[j + 12]));
#pragma unroll
        for (uint32_t b = 0; b < 16; b += 4) s[b + j] = madd(s[b + j], sum);
    }
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) s[i] = pos_pow5(s[i]);
}
__global__ __launch_bounds__(256) void poseidon_trace_kernel(uint32_t log_n, uint32_t* __restrict__ out,
                                                             uint64_t stride) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t row = t & ((1u << log_n) - 1), rep = t >> log_n;
    if (rep >= 8) return;
    uint32_t* o = out + (uint64_t)rep * POS_COLS_PER_REP * stride + row;
    uint32_t s[16];
#pragma unroll
    for (uint32_t i = 0; i < 16; i++) {
        s[i] = (uint32_t)(((uint64_t)row * 16 + i + rep) % P);
        o[(uint64_t)i * stride] = s[i];
    }
    uint32_t col = 16;
    for (uint32_t r = 0; r < 4; r++) {
        pos_external(s, r);
#pragma unroll
        for (uint32_t i = 0; i < 16; i++) o[(uint64_t)(col + i) * stride] = s[i];
        col += 16;
    }
    for (uint32_t r = 0; r < 14; r++) {
        // internal round (input.zig:135-139,203-210): state[0] += c, state[i] = state[i]*2^(i+1) + sum
        s[0] = madd(s[0], 9876u + r * 17u);
        uint32_t sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < 16; i++) sum = madd(sum, s[i]);
#pragma unroll
        for (uint32_t i = 0; i < 16; i++) s[i] = madd(mmul(s[i], 2u << i), sum);
        s[0] = pos_pow5(s[0]);
        o[(uint64_t)col * stride] = s[0];
        col++;
    }
    for (uint32_t r = 4; r < 8; r++) {
        pos_external(s, r);
#pragma unroll
        for (uint32_t i = 0; i < 16; i++) o[(uint64_t)(col + i) * stride] = s[i];
        col += 16;
    }
}
void launch_poseidon_trace(Ctx& c, uint32_t log_n, uint32_t* d_out, uint64_t col_stride) {
    SR_CHECK(log_n >= 1 && log_n <= 27, "poseidon trace: log_n_rows out of range");
    const uint64_t threads = 8ull << log_n;
    KTimer kt(c, "poseidon_trace", 4.0 * 8 * POS_COLS_PER_REP * (double)(1ull << log_n));
    hipLaunchKernelGGL(poseidon_trace_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, c.stream, log_n,
                       d_out, col_stride);
    SR_HIP(hipGetLastError());
}

// Periodic XOR example AIR columns (examples/xor/input.zig:31-107; core/air/utils.zig:23-65):
// thread i takes coset row i to its bit-reversed circle-domain position (a permutation, so
// every word of main and is_step is written once) and writes storage word i of is_first.
__global__ __launch_bounds__(256) void xor_example_columns_kernel(uint32_t log_size, uint32_t step_mask,
                                                                  uint32_t offset_mod, uint32_t* __restrict__ main_col,
                                                                  uint32_t* __restrict__ is_first,
                                                                  uint32_t* __restrict__ is_step) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = 1u << log_size;
    if (i >= n) return;
    const uint32_t ci = (i & 1) == 0 ? i / 2 : (2 * n - i) / 2;
    const uint32_t pos = bitrev32(ci, log_size);
    main_col[pos] = (i & 1) == 0 ? 1u : 0u;
    is_step[pos] = (i & step_mask) == offset_mod ? 1u : 0u;
    is_first[i] = i == 0 ? 1u : 0u;
}
void launch_xor_example_columns(Ctx& c, uint32_t log_size, uint32_t log_step, uint64_t offset, uint32_t* d_main,
                                uint32_t* d_is_first, uint32_t* d_is_step) {
    SR_CHECK(log_size >= 1 && log_size <= 30 && log_step <= log_size, "xor columns: log sizes");
    const uint32_t n = 1u << log_size;
    const uint32_t mask = (uint32_t)((1ull << log_step) - 1);
    hipLaunchKernelGGL(xor_example_columns_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, log_size, mask,
                       (uint32_t)(offset & mask), d_main, d_is_first, d_is_step);
    SR_HIP(hipGetLastError());
}

// A constant composition evaluation (accumulation.zig:142-159 with one constraint, whose
// coefficient is 1): every row of the four coordinate columns holds v.
__global__ __launch_bounds__(256) void fill_qm31_kernel(uint32_t n, uint4 v, uint32_t* __restrict__ o0,
                                                        uint32_t* __restrict__ o1, uint32_t* __restrict__ o2,
                                                        uint32_t* __restrict__ o3) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    o0[i] = v.x;
    o1[i] = v.y;
    o2[i] = v.z;
    o3[i] = v.w;
}
void launch_constant_composition(Ctx& c, uint32_t log_eval, const uint32_t v[4], uint32_t* const* out) {
    const uint32_t n = 1u << log_eval;
    hipLaunchKernelGGL(fill_qm31_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, n,
                       make_uint4(v[0], v[1], v[2], v[3]), out[0], out[1], out[2], out[3]);
    SR_HIP(hipGetLastError());
}

// Point evaluation (prover/poly/circle/point_evaluation.zig:14-38,469-488):
// p(P) = sum_i c_i * w_i, w_i = prod_{bit j of i} f_j, f = [P.y, P.x, 2x^2-1, ...].
struct Factors {
    qm31 f[30];
};
// w_i = T_hi[i >> h] * T_lo[i & (2^h - 1)] with T_lo[l] = prod_{bit j of l} f_j (j < h) and
// T_hi[u] = prod_{bit j of u} f_{h+j}: one product per weight (field products are exact, so
// the split gives the same values).  The tables come first, one thread per entry.
__global__ __launch_bounds__(256) void eval_weight_tables_kernel(Factors fs, uint32_t log_size, uint32_t h,
                                                                 uint4* __restrict__ tab) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_lo = 1u << h, n_hi = 1u << (log_size - h);
    if (k >= n_lo + n_hi) return;
    const bool lo = k < n_lo;
    const uint32_t idx = lo ? k : k - n_lo, j0 = lo ? 0 : h, nb = lo ? h : log_size - h;
    qm31 acc = qone();
    for (uint32_t j = 0; j < nb; j++)
        if ((idx >> j) & 1) acc = qmul(acc, fs.f[j0 + j]);
    tab[k] = make_uint4(acc.c0.a, acc.c0.b, acc.c1.a, acc.c1.b);
}
__global__ __launch_bounds__(256) void eval_weight_tables_dev_kernel(const uint4* __restrict__ fs, uint32_t log_size,
                                                                     uint32_t h, uint4* __restrict__ tab,
                                                                     uint64_t bstride) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    fs = boffp(fs, boff);
    tab = boffp(tab, boff);
    const uint32_t n_lo = 1u << h, n_hi = 1u << (log_size - h);
    if (k >= n_lo + n_hi) return;
    const bool lo = k < n_lo;
    const uint32_t idx = lo ? k : k - n_lo, j0 = lo ? 0 : h, nb = lo ? h : log_size - h;
    qm31 acc = qone();
    for (uint32_t j = 0; j < nb; j++)
        if ((idx >> j) & 1) {
            const uint4 f = fs[j0 + j];
            acc = qmul(acc, qm31{{f.x, f.y}, {f.z, f.w}});
        }
    tab[k] = make_uint4(acc.c0.a, acc.c0.b, acc.c1.a, acc.c1.b);
}
__global__ __launch_bounds__(256) void eval_weights_kernel(const uint4* __restrict__ tab, uint32_t log_size, uint32_t h,
                                                           uint4* __restrict__ w, uint64_t bstride) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (1u << log_size)) return;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    tab = boffp(tab, boff);
    w = boffp(w, boff);
    const uint4 a = tab[i & ((1u << h) - 1)], b = tab[(1u << h) + (i >> h)];
    const qm31 v = qmul(qm31{{a.x, a.y}, {a.z, a.w}}, qm31{{b.x, b.y}, {b.z, b.w}});
    w[i] = make_uint4(v.c0.a, v.c0.b, v.c1.a, v.c1.b);
}
uint64_t eval_table_bytes(uint32_t log_size) {
    const uint32_t h = log_size / 2;
    return 16ull * ((1ull << h) + (1ull << (log_size - h)));
}

void launch_eval_weights_into(Ctx& c, const qm31 factors[32], uint32_t log_size, uint32_t* d_w, uint32_t* d_tab) {
    SR_CHECK(log_size <= 30, "eval weights: log size too large");
    Factors fs;
    for (uint32_t j = 0; j < log_size; j++) fs.f[j] = factors[j];
    const uint32_t n = 1u << log_size, h = log_size / 2;
    const uint32_t nt = (1u << h) + (1u << (log_size - h));
    KTimer kt(c, "eval_weights", 16.0 * n);
    hipLaunchKernelGGL(eval_weight_tables_kernel, dim3((nt + 255) / 256), dim3(256), 0, c.stream, fs, log_size, h,
                       (uint4*)d_tab);
    hipLaunchKernelGGL(eval_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, (const uint4*)d_tab,
                       log_size, h, (uint4*)d_w, (uint64_t)0);
    SR_HIP(hipGetLastError());
}

void launch_eval_weights(Ctx& c, const qm31 factors[32], uint32_t log_size, uint32_t* d_w) {
    SR_CHECK(log_size <= 30, "eval weights: log size too large");
    DBuf tab(&c, eval_table_bytes(log_size));
    launch_eval_weights_into(c, factors, log_size, d_w, tab.as<uint32_t>());
}

void launch_eval_weights_dev(Ctx& c, const uint32_t* d_factors, uint32_t log_size, uint32_t* d_w, uint32_t batch,
                             uint64_t bstride, uint32_t* d_tab) {
    SR_CHECK(log_size <= 30, "eval weights: log size too large");
    SR_CHECK(batch >= 1 && batch <= 65535 && (batch == 1 || d_tab), "eval weights: batch");
    const uint32_t n = 1u << log_size, h = log_size / 2;
    const uint32_t nt = (1u << h) + (1u << (log_size - h));
    DBuf tab_own;
    if (!d_tab) {
        tab_own = DBuf(&c, 16ull * nt);
        d_tab = tab_own.as<uint32_t>();
    }
    KTimer kt(c, "eval_weights", batch * 16.0 * n);
    hipLaunchKernelGGL(eval_weight_tables_dev_kernel, dim3((nt + 255) / 256, batch), dim3(256), 0, c.stream,
                       (const uint4*)d_factors, log_size, h, (uint4*)d_tab, bstride);
    hipLaunchKernelGGL(eval_weights_kernel, dim3((n + 255) / 256, batch), dim3(256), 0, c.stream, (const uint4*)d_tab,
                       log_size, h, (uint4*)d_w, bstride);
    SR_HIP(hipGetLastError());
}

// One block per (chunk, column group of EV_CB): each block loads a weight once for EV_CB
// columns, accumulates raw 62-bit products four elements at a time (< 2^64) before one
// Mersenne fold, and stores its 4 * EV_CB partial sums in its own slot; a second kernel adds
// the slots (integer sums are exact and order independent).  Workgroups are dealt to the
// 8 XCDs round-robin, so block b runs on XCD b % 8: the blocks b, b + 8, ... of one XCD
// take all column groups of one chunk in a row, and the chunk's weights are read from HBM
// once and then served by that XCD's L2.
constexpr uint32_t EV_PER_THREAD = 16;
constexpr uint32_t EV_CB = 4;
constexpr uint32_t EV_XCDS = 8;
__device__ __forceinline__ uint64_t ev_wave_sum(uint64_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m), hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}
__global__ __launch_bounds__(256) void eval_columns_kernel(const uint32_t* const* __restrict__ cols, uint32_t n_cols,
                                                           uint32_t log_size, const uint4* __restrict__ w,
                                                           uint32_t chunks, uint32_t groups,
                                                           unsigned long long* __restrict__ part, uint64_t bstride) {
    __shared__ unsigned long long red[4][EV_CB * 4];
    const uint32_t t = threadIdx.x;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;  // proof blockIdx.y of a batch
    w = boffp(w, boff);
    part = boffp(part, boff);
    const uint32_t per = chunks < EV_XCDS ? chunks : EV_XCDS;  // fewer chunks than XCDs: no idle blocks
    const uint32_t k = blockIdx.x / per;
    const uint32_t chunk = (k / groups) * per + blockIdx.x % per;
    if (chunk >= chunks) return;
    const uint32_t c0 = (k % groups) * EV_CB;
    const uint32_t nc = n_cols - c0 < EV_CB ? n_cols - c0 : EV_CB;
    const uint32_t n = 1u << log_size;
    uint64_t s[EV_CB][4];
#pragma unroll
    for (uint32_t c = 0; c < EV_CB; c++) s[c][0] = s[c][1] = s[c][2] = s[c][3] = 0;
    if (n >= 4) {
        // lane t takes 4 consecutive elements per group: one 16-byte coefficient load per column
        // and 64 contiguous bytes of weights (the block covers 256 * EV_PER_THREAD elements)
#pragma unroll
        for (uint32_t g = 0; g < EV_PER_THREAD / 4; g++) {
            const uint32_t i0 = chunk * (256 * EV_PER_THREAD) + g * 1024 + 4 * t;
            if (i0 >= n) break;
            uint4 ww[4];
#pragma unroll
            for (int e = 0; e < 4; e++) ww[e] = w[i0 + e];
            // every column load is issued before the first is used (a short last group
            // re-reads its last column instead of branching around the loads)
            u32x4 cv[EV_CB];
#pragma unroll
            for (uint32_t c = 0; c < EV_CB; c++)
                cv[c] = *(const SR_GLOBAL u32x4*)((gcu32*)boffp(cols[c0 + (c < nc ? c : nc - 1)], boff) + i0);
#pragma unroll
            for (uint32_t c = 0; c < EV_CB; c++) {
                if (c < nc) {
                    const u32x4 v4 = cv[c];
                    const uint64_t v[4] = {v4.x, v4.y, v4.z, v4.w};
                    uint64_t r0 = 0, r1 = 0, r2 = 0, r3 = 0;
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        r0 += v[e] * ww[e].x;
                        r1 += v[e] * ww[e].y;
                        r2 += v[e] * ww[e].z;
                        r3 += v[e] * ww[e].w;
                    }
                    s[c][0] += mfold(r0);
                    s[c][1] += mfold(r1);
                    s[c][2] += mfold(r2);
                    s[c][3] += mfold(r3);
                }
            }
        }
    } else if (t < n) {
        const uint4 ww = w[t];
#pragma unroll
        for (uint32_t c = 0; c < EV_CB; c++) {
            if (c < nc) {
                const uint64_t v = boffp(cols[c0 + c], boff)[t];
                s[c][0] += mfold(v * ww.x);
                s[c][1] += mfold(v * ww.y);
                s[c][2] += mfold(v * ww.z);
                s[c][3] += mfold(v * ww.w);
            }
        }
    }
#pragma unroll
    for (uint32_t c = 0; c < EV_CB; c++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t x = ev_wave_sum(s[c][k]);
            if ((t & 63) == 0) red[t >> 6][4 * c + k] = x;
        }
    __syncthreads();
    if (t < 4 * nc) {
        const uint64_t x = red[0][t] + red[1][t] + red[2][t] + red[3][t];
        part[(uint64_t)(4 * c0 + t) * chunks + chunk] = x;
    }
}
// out[k] = sum over chunks of part[k][chunk] mod p: one wave per output word.
__global__ __launch_bounds__(256) void eval_reduce_kernel(const unsigned long long* __restrict__ part, uint32_t chunks,
                                                          uint32_t n_out, uint32_t* __restrict__ out, uint64_t bstride) {
    const uint32_t k = blockIdx.x * 4 + (threadIdx.x >> 6), l = threadIdx.x & 63;
    if (k >= n_out) return;
    const uint64_t boff = (uint64_t)blockIdx.y * bstride;
    part = boffp(part, boff);
    out = boffp(out, boff);
    uint64_t x = 0;
    for (uint32_t j = l; j < chunks; j += 64) x += part[(uint64_t)k * chunks + j];
    x = ev_wave_sum(x);
    if (l == 0) out[k] = mred64(x);
}
static uint32_t eval_chunks(uint32_t log_size) {
    return ((1u << log_size) + 256 * EV_PER_THREAD - 1) / (256 * EV_PER_THREAD);
}

uint64_t eval_partial_bytes(uint32_t n_cols, uint32_t log_size) {
    return 8ull * 4 * n_cols * eval_chunks(log_size);
}

void launch_eval_columns_into(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_size,
                              const uint32_t* d_w, uint32_t* d_out, unsigned long long* d_part, uint32_t batch,
                              uint64_t bstride) {
    if (n_cols == 0) return;
    SR_CHECK(batch >= 1 && batch <= 65535, "eval columns: batch size");
    const uint32_t n = 1u << log_size;
    const uint32_t chunks = eval_chunks(log_size);
    const uint32_t groups = (n_cols + EV_CB - 1) / EV_CB;
    const uint32_t nw = n_cols * 4;
    KTimer kt(c, "eval_columns", batch * (4.0 * n_cols + 16.0) * n);
    const uint32_t per = chunks < EV_XCDS ? chunks : EV_XCDS;
    const uint32_t blocks = (chunks + per - 1) / per * per * groups;
    hipLaunchKernelGGL(eval_columns_kernel, dim3(blocks, batch), dim3(256), 0, c.stream, d_cols, n_cols, log_size,
                       (const uint4*)d_w, chunks, groups, d_part, bstride);
    hipLaunchKernelGGL(eval_reduce_kernel, dim3((nw + 3) / 4, batch), dim3(256), 0, c.stream,
                       (const unsigned long long*)d_part, chunks, nw, d_out, bstride);
    SR_HIP(hipGetLastError());
}

void launch_eval_columns(Ctx& c, const uint32_t* const* d_cols, uint32_t n_cols, uint32_t log_size,
                         const uint32_t* d_w, uint32_t* d_out) {
    if (n_cols == 0) return;
    DBuf part(&c, eval_partial_bytes(n_cols, log_size));
    launch_eval_columns_into(c, d_cols, n_cols, log_size, d_w, d_out, part.as<unsigned long long>());
}

}  // namespace sr
