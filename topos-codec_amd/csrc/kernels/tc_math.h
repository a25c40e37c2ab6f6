/* Integer helpers shared by the C11 host layer and the gfx950 kernels.
 * Compiled as plain C on the host (TC_HD empty) and as __host__ __device__
 * inline functions inside the HIP translation units, so the CPU test suite
 * exercises exactly the arithmetic the kernels run.
 */
#ifndef TC_MATH_H
#define TC_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define TC_HD __host__ __device__ static inline
#else
#define TC_HD static inline
#endif

#define TC_CRC_POLY 0xEDB88320u

/* a * b mod P over GF(2) in the reflected bit order of zlib's CRC32
 * (bit 31 is x^0). Fixed 32-step loop: no data-dependent trip count. */
TC_HD uint32_t tc_gf2_mulmod(uint32_t a, uint32_t b)
{
    uint32_t p = 0u;
    for (int i = 0; i < 32; ++i) {
        if (a & (0x80000000u >> i)) { p ^= b; }
        b = (b & 1u) ? (b >> 1) ^ TC_CRC_POLY : (b >> 1);
    }
    return p;
}

/* x^(8*nbytes) mod P, from a table pw[k] = x^(2^k) mod P (k = 0..31). */
TC_HD uint32_t tc_gf2_xpow8n(const uint32_t* pw, uint64_t nbytes)
{
    uint32_t p = 0x80000000u; /* x^0 */
    uint32_t k = 3u;          /* 8 = 2^3 bits per byte */
    while (nbytes != 0u) {
        if (nbytes & 1u) { p = tc_gf2_mulmod(pw[k & 31u], p); }
        nbytes >>= 1;
        ++k;
    }
    return p;
}

/* Shift a raw (zero-initialised, un-inverted) CRC register over nbytes of
 * zero data: the term a prefix contributes after nbytes more input. */
TC_HD uint32_t tc_crc_shift(const uint32_t* pw, uint32_t crc, uint64_t nbytes)
{
    return tc_gf2_mulmod(tc_gf2_xpow8n(pw, nbytes), crc);
}

/* Exact floor(x / f) for x < 2^31, 1 <= f <= 4096 (the rANS state update):
 * with l = ceil(log2 f) and m = ceil(2^(31+l) / f) < 2^32,
 * floor(x/f) = umulhi(x, m) >> (l - 1) because x * (m*f - 2^(31+l)) < 2^(31+l).
 * f == 1 is handled by the caller (q = x). */
TC_HD void tc_rans_recip(uint32_t f, uint32_t* mag, uint32_t* sh)
{
    if (f <= 1u) { *mag = 0u; *sh = 0u; return; }
    uint32_t l = 0u;
    while ((1u << l) < f) { ++l; }
    const uint64_t num = (uint64_t)1 << (31u + l);
    *mag = (uint32_t)((num + f - 1u) / f);
    *sh = l - 1u;
}

/* Encoder constants of one symbol with frequency f (1..4096) and cumulative
 * frequency cum, for the update x' = (x / f) * 4096 + x % f + cum:
 *   A = ceil(2^(31+l) / f) with l = ceil(log2 f); since x < f * 2^19 <= 2^31
 *       after renormalisation, floor(x / f) = umulhi(2x, A) >> l exactly
 *       (A * f - 2^(31+l) < f <= 2^l, so 2x * (A*f - 2^(31+l)) < 2^(32+l));
 *       f = 1 gives A = 2^31, l = 0, i.e. x itself;
 *   B = cum | (4096 - f) << 12 | l << 24, so x' = x + q * (4096 - f) + cum;
 *   X = f << 19, the bound the state must be below before the update. */
TC_HD void tc_rans_enc_consts(uint32_t f, uint32_t cum, uint32_t* A, uint32_t* B, uint32_t* X)
{
    uint32_t l = 0u;
    while ((1u << l) < f) { ++l; }
    const uint64_t num = (uint64_t)1 << (31u + l);
    *A = f == 0u ? 0u : (uint32_t)((num + f - 1u) / f);
    *B = (cum & 0xFFFu) | (((4096u - f) & 0xFFFu) << 12) | (l << 24);
    *X = f << 19;
}

TC_HD uint32_t tc_rans_enc_update(uint32_t x, uint32_t A, uint32_t B)
{
    const uint32_t q = (uint32_t)(((uint64_t)(x << 1) * A) >> 32) >> (B >> 24);
    return x + q * ((B >> 12) & 0xFFFu) + (B & 0xFFFu);
}

/* Exact floor(n / Q) for n < 2^31, Q >= 1, with one 32-bit high multiply
 * (division by an invariant integer, Granlund and Montgomery 1994, thm 4.2):
 * with l = ceil(log2 Q) and m = ceil(2^(31+l) / Q), 0 <= m Q - 2^(31+l) < Q
 * <= 2^l, hence floor(n / Q) = floor(m n / 2^(31+l)) for every n < 2^31.
 * m < 2^32 (m = 2^31 for Q a power of two, m < 2^32 otherwise since
 * Q > 2^(l-1)), and m n / 2^(31+l) = umulhi(2n, m) / 2^l with 2n < 2^32.
 * (the quantiser's divide on the GPU) */
TC_HD void tc_qdiv32_consts(uint32_t Q, uint32_t* m, uint32_t* l)
{
    uint32_t s = 0u;
    while (((uint64_t)1 << s) < Q) { ++s; }
    *m = (uint32_t)((((uint64_t)1 << (31u + s)) + Q - 1u) / Q);
    *l = s;
}

TC_HD uint32_t tc_qdiv32(uint32_t n, uint32_t m, uint32_t l)
{
    return (uint32_t)(((uint64_t)(n << 1) * m) >> 32) >> l;
}

/* Exact floor(n / d) for n * d < 2^51 via m = floor(2^51 / d) + 1:
 * floor(n / d) = (n * m) >> 51 (the quantiser's divide, bit depth <= 12). */
TC_HD uint64_t tc_qdiv_magic(uint32_t d) { return (((uint64_t)1 << 51) / d) + 1u; }

TC_HD uint32_t tc_qdiv_apply(uint32_t n, uint64_t m)
{
    const uint64_t lo = (uint64_t)n * (uint32_t)m;
    const uint64_t hi = (uint64_t)n * (uint32_t)(m >> 32);
    return (uint32_t)(((lo >> 32) + hi) >> 19);
}

/* ---- rANS tables ---- */

/* 8-bit proportional row bytes: round-half-up share of 255, used symbols at
 * least 1, an empty row becomes {1, 0, ...} (the per-context table format of
 * the V7-R2 slice prefix). */
TC_HD void tc_row_bytes(uint8_t* out, uint32_t nsym, const uint32_t* count)
{
    uint64_t total = 0u;
    for (uint32_t s = 0; s < nsym; ++s) { total += count[s]; }
    if (total == 0u) {
        for (uint32_t s = 0; s < nsym; ++s) { out[s] = 0u; }
        out[0] = 1u;
        return;
    }
    for (uint32_t s = 0; s < nsym; ++s) {
        if (count[s] == 0u) { out[s] = 0u; continue; }
        const uint64_t num = (uint64_t)count[s] * 255u + total / 2u;
        uint64_t v;
        if ((num >> 32) == 0u && (total >> 32) == 0u) { v = (uint32_t)num / (uint32_t)total; }
        else { v = num / total; }
        if (v == 0u) { v = 1u; }
        if (v > 255u) { v = 255u; }
        out[s] = (uint8_t)v;
    }
}

/* Normalise byte counts (<= 255 each) to frequencies summing to 4096:
 * floor share with a floor of 1 for used symbols, then remove one from the
 * largest frequency (first on ties) while the sum is over, then add one to
 * the largest remainder (first on ties) while under. Returns 0 on success,
 * -1 for an all-zero row. Cum is written with cum[nsym] = 4096. */
TC_HD int tc_model_from_bytes(const uint8_t* bytes, uint32_t nsym, uint16_t* freq, uint16_t* cum)
{
    uint32_t total = 0u;
    for (uint32_t s = 0; s < nsym; ++s) { total += bytes[s]; }
    if (total == 0u) { return -1; }
    uint32_t f[64], rem[64], sum = 0u;
    for (uint32_t s = 0; s < nsym; ++s) {
        const uint32_t sc = (uint32_t)bytes[s] * 4096u;
        f[s] = bytes[s] == 0u ? 0u : sc / total;
        rem[s] = sc % total;
        if (bytes[s] != 0u && f[s] == 0u) { f[s] = 1u; }
        sum += f[s];
    }
    while (sum > 4096u) {
        uint32_t big = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (f[s] > 1u && (big == 0xFFFFFFFFu || f[s] > f[big])) { big = s; }
        }
        if (big == 0xFFFFFFFFu) { break; }
        f[big] -= 1u; rem[big] = 0u; sum -= 1u;
    }
    while (sum < 4096u) {
        uint32_t best = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (bytes[s] == 0u) { continue; }
            if (best == 0xFFFFFFFFu || rem[s] > rem[best]) { best = s; }
        }
        if (best == 0xFFFFFFFFu) { break; }
        f[best] += 1u; rem[best] = 0u; sum += 1u;
    }
    if (sum != 4096u) { return -1; }
    uint32_t acc = 0u;
    for (uint32_t s = 0; s < nsym; ++s) {
        freq[s] = (uint16_t)f[s];
        cum[s] = (uint16_t)acc;
        acc += f[s];
    }
    cum[nsym] = (uint16_t)acc;
    return 0;
}

TC_HD uint32_t tc_bitlen(uint32_t m)
{
    uint32_t n = 0u;
    while (m != 0u) { ++n; m >>= 1; }
    return n;
}

TC_HD uint32_t tc_map_signed(int32_t v)
{
    const int64_t s = v;
    return s >= 0 ? (uint32_t)((uint64_t)s << 1) : (uint32_t)(((uint64_t)(-s) << 1) - 1u);
}

TC_HD int32_t tc_unmap_signed(uint32_t m)
{
    return (m & 1u) == 0u ? (int32_t)(m >> 1) : -(int32_t)(((uint64_t)m + 1u) >> 1);
}

/* 8-point transform with the integer matrix M of the codec (rows: 13 13 13 13
 * 13 13 13 13 / 13 11 7 1 -1 -7 -11 -13 / 12 5 -5 -12 -12 -5 5 12 / 11 -4 -13
 * -8 8 13 4 -11 / 9 -9 -9 9 9 -9 -9 9 / 8 -13 4 11 -11 -4 13 -8 / 5 -12 12 -5
 * -5 12 -12 5 / 1 -7 11 -13 13 -11 7 -1): even rows are symmetric and odd rows
 * antisymmetric, so out = M * in splits into sums s_k = in_k + in_{7-k} (even
 * rows) and differences d_k = in_k - in_{7-k} (odd rows); rows 0/4 and 2/6
 * split once more. Every step is an exact integer operation, so the result is
 * identical to the direct product (36 operations instead of 64).
 * Forward: out[v] = sum_k M[v][k] in[k]. */
TC_HD void tc_fdct8_i32(const int32_t* in, int32_t* out)
{
    const int32_t s0 = in[0] + in[7], s1 = in[1] + in[6], s2 = in[2] + in[5], s3 = in[3] + in[4];
    const int32_t d0 = in[0] - in[7], d1 = in[1] - in[6], d2 = in[2] - in[5], d3 = in[3] - in[4];
    const int32_t e0 = s0 + s3, e1 = s1 + s2, f0 = s0 - s3, f1 = s1 - s2;
    out[0] = 13 * (e0 + e1);
    out[4] = 9 * (e0 - e1);
    out[2] = 12 * f0 + 5 * f1;
    out[6] = 5 * f0 - 12 * f1;
    out[1] = 13 * d0 + 11 * d1 + 7 * d2 + 1 * d3;
    out[3] = 11 * d0 - 4 * d1 - 13 * d2 - 8 * d3;
    out[5] = 8 * d0 - 13 * d1 + 4 * d2 + 11 * d3;
    out[7] = 1 * d0 - 7 * d1 + 11 * d2 - 13 * d3;
}

/* Inverse direction: out[x] = sum_v in[v] M[v][x] (the transpose), with
 * out[x] = E_x + O_x and out[7-x] = E_x - O_x. T is double (all values are
 * integers below 2^53, so every sum is exact) or int64_t. */
#define TC_IDCT8_BODY(T)                                                          \
    const T p = (T)13 * in[0] + (T)9 * in[4], q = (T)13 * in[0] - (T)9 * in[4];    \
    const T r = (T)12 * in[2] + (T)5 * in[6], t = (T)5 * in[2] - (T)12 * in[6];    \
    const T e0 = p + r, e3 = p - r, e1 = q + t, e2 = q - t;                         \
    const T o0 = (T)13 * in[1] + (T)11 * in[3] + (T)8 * in[5] + (T)1 * in[7];       \
    const T o1 = (T)11 * in[1] - (T)4 * in[3] - (T)13 * in[5] - (T)7 * in[7];       \
    const T o2 = (T)7 * in[1] - (T)13 * in[3] + (T)4 * in[5] + (T)11 * in[7];       \
    const T o3 = (T)1 * in[1] - (T)8 * in[3] + (T)11 * in[5] - (T)13 * in[7];       \
    out[0] = e0 + o0; out[7] = e0 - o0;                                             \
    out[1] = e1 + o1; out[6] = e1 - o1;                                             \
    out[2] = e2 + o2; out[5] = e2 - o2;                                             \
    out[3] = e3 + o3; out[4] = e3 - o3;
TC_HD void tc_idct8_f64(const double* in, double* out) { TC_IDCT8_BODY(double) }
TC_HD void tc_idct8_i64(const int64_t* in, int64_t* out) { TC_IDCT8_BODY(int64_t) }

/* Scan (zigzag) position of natural coefficient index n = 8r + c in the
 * 8x8 zigzag order that starts along the first row: diagonal d = r + c holds
 * positions from d(d+1)/2 (d < 8) or 64 - (15-d)(16-d)/2 (d >= 8); even
 * diagonals run bottom-left to top-right, odd ones top-right to bottom-left. */
TC_HD uint32_t tc_zigzag_pos(uint32_t n)
{
    const uint32_t r = n >> 3, c = n & 7u, d = r + c;
    if (d < 8u) { return d * (d + 1u) / 2u + ((d & 1u) ? r : c); }
    return 64u - (15u - d) * (16u - d) / 2u + ((d & 1u) ? 7u - c : 7u - r);
}

/* context buckets of the V7-R2 conditional models */
TC_HD uint32_t tc_ctx_pos(uint32_t pos) { return pos < 8u ? 0u : pos < 16u ? 1u : pos < 32u ? 2u : 3u; }
TC_HD uint32_t tc_ctx_prev(uint32_t cat, uint32_t has_prev)
{
    return has_prev == 0u ? 4u : cat == 0u ? 0u : cat == 1u ? 1u : cat <= 3u ? 2u : 3u;
}
TC_HD uint32_t tc_ctx_dc(uint32_t cat, uint32_t has_prev)
{
    return has_prev == 0u ? 0u : 1u + (cat == 0u ? 0u : cat == 1u ? 1u : cat <= 3u ? 2u : 3u);
}

/* The same buckets as packed-nibble lookups (the decode kernel's form):
 * position buckets by idx >> 3 -> (0, 1, 2, 2, 3, 3, 3, 3); previous-category
 * buckets from the decoder's entry fields w = max(cat - 1, 0), nz = (cat != 0):
 * cat = w + nz, 0 for cat 0, else by min(w, 7) -> (1, 2, 2, 3, 3, 3, 3, 3). */
#define TC_CTX_POS_NIBBLES 0x33332210u
#define TC_CTX_PREV_NIBBLES 0x33333221u
TC_HD uint32_t tc_ctx_pos_nib(uint32_t idx) { return (TC_CTX_POS_NIBBLES >> ((idx >> 1) & 0x1Cu)) & 0xFu; }
TC_HD uint32_t tc_ctx_prev_wnz(uint32_t w, uint32_t nz)
{
    return nz ? (TC_CTX_PREV_NIBBLES >> ((w < 7u ? w : 7u) * 4u)) & 0xFu : 0u;
}

TC_HD int32_t tc_dc_predict(uint32_t has_left, int32_t left, uint32_t has_top, int32_t top)
{
    if (has_left && has_top) {
        const int64_t s = (int64_t)left + top;
        return s >= 0 ? (int32_t)((s + 1) >> 1) : -(int32_t)((-s) >> 1);
    }
    if (has_left) { return left; }
    if (has_top) { return top; }
    return 0;
}

/* Inverse-transform weights W[u][v] = round(2^32 / (E_u E_v)) of the codec,
 * row-major (native/src/transform/transform_tables.h:10-36): a bitstream
 * constant, shared by the device table and the host's dequantisation weights. */
#define TC_TW_INIT                                                                                   \
    {                                                                                                \
        2350, 4672, 4699, 4293, 4902, 4293, 4699, 4672,  4672, 9288, 9343, 8535, 9747, 8535, 9343, 9288, \
        4699, 9343, 9399, 8586, 9805, 8586, 9399, 9343,  4293, 8535, 8586, 7843, 8957, 7843, 8586, 8535, \
        4902, 9747, 9805, 8957, 10228, 8957, 9805, 9747, 4293, 8535, 8586, 7843, 8957, 7843, 8586, 8535, \
        4699, 9343, 9399, 8586, 9805, 8586, 9399, 9343,  4672, 9288, 9343, 8535, 9747, 8535, 9343, 9288, \
    }

#endif
