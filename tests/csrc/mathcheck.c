/* Exposes the host/device shared integer helpers (tc_math.h) to the CPU test
 * suite so the exact arithmetic the kernels run can be checked on the host. */
#include <stdint.h>

#include "tc_math.h"

uint32_t mc_gf2_mulmod(uint32_t a, uint32_t b) { return tc_gf2_mulmod(a, b); }
uint32_t mc_crc_shift(const uint32_t* pw, uint32_t crc, uint64_t n) { return tc_crc_shift(pw, crc, n); }
void mc_rans_recip(uint32_t f, uint32_t* mag, uint32_t* sh) { tc_rans_recip(f, mag, sh); }
uint64_t mc_qdiv_magic(uint32_t d) { return tc_qdiv_magic(d); }
uint32_t mc_qdiv_apply(uint32_t n, uint64_t m) { return tc_qdiv_apply(n, m); }
void mc_row_bytes(uint8_t* out, uint32_t nsym, const uint32_t* count) { tc_row_bytes(out, nsym, count); }
int mc_model(const uint8_t* b, uint32_t nsym, uint16_t* f, uint16_t* c) { return tc_model_from_bytes(b, nsym, f, c); }

/* rANS state division as the kernels compute it */
uint32_t mc_rans_div(uint32_t x, uint32_t f)
{
    uint32_t mag, sh;
    tc_rans_recip(f, &mag, &sh);
    if (f == 1u) { return x; }
    return (uint32_t)(((uint64_t)x * mag) >> 32) >> sh;
}

/* wave-style chunked CRC32 (64 chunks, shifted and xor-combined) */
uint32_t mc_crc_chunked(const uint8_t* p, uint32_t n, const uint32_t* tab, const uint32_t* pw, uint32_t parts)
{
    const uint32_t per = (n + parts - 1u) / parts;
    uint32_t acc = 0u;
    for (uint32_t l = 0; l < parts; ++l) {
        const uint32_t a = l * per < n ? l * per : n;
        const uint32_t b = a + per < n ? a + per : n;
        uint32_t c = 0u;
        for (uint32_t i = a; i < b; ++i) { c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8); }
        if (a < b) { acc ^= tc_crc_shift(pw, c, (uint64_t)(n - b)); }
    }
    return acc ^ tc_crc_shift(pw, 0xFFFFFFFFu, n) ^ 0xFFFFFFFFu;
}

/* encoder update with precomputed constants against the plain division, for
 * every frequency 1..4096: state edges around every multiple boundary of f
 * near the renormalisation bound plus pseudo-random states below it. Returns
 * the number of mismatches. */
uint32_t mc_enc_update_check(uint32_t per_f)
{
    uint32_t bad = 0u;
    uint64_t s = 0x9E3779B97F4A7C15ull;
    for (uint32_t f = 1u; f <= 4096u; ++f) {
        uint32_t A, B, X;
        const uint32_t cum = (f * 2654435761u) % (4097u - f);
        tc_rans_enc_consts(f, cum, &A, &B, &X);
        if (X != (f << 19)) { ++bad; }
        const uint32_t hi = X - 1u < 0x7FFFFFFFu ? X - 1u : 0x7FFFFFFFu;
        for (uint32_t i = 0; i < per_f + 8u; ++i) {
            uint32_t x;
            if (i < 4u) { x = hi - i; }
            else if (i < 8u) { x = (uint32_t)((uint64_t)hi / f * f) - 2u + i - 4u; }
            else {
                s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                x = (uint32_t)(s % ((uint64_t)hi + 1u));
            }
            if (x > hi) { continue; }
            const uint64_t want = (uint64_t)(x / f) * 4096u + x % f + cum;
            if ((uint64_t)tc_rans_enc_update(x, A, B) != want) { ++bad; }
        }
    }
    return bad;
}

uint32_t mc_zigzag_pos(uint32_t n) { return tc_zigzag_pos(n); }

/* the quantiser's 32-bit multiply-high divide against integer division:
 * every Q in qs[0..nq), numerators around every multiple boundary near the
 * top of the range (n < 2^31), small numerators and pseudo-random ones, and
 * the constants' ranges (m < 2^32 by construction of the 64-bit division
 * below, l <= 31). Returns mismatches. */
uint32_t mc_qdiv32_check(const uint32_t* qs, uint32_t nq, uint32_t per_q)
{
    uint32_t bad = 0u;
    uint64_t s = 0x2545F4914F6CDD1Dull;
    for (uint32_t i = 0; i < nq; ++i) {
        const uint32_t Q = qs[i];
        uint32_t m, l;
        tc_qdiv32_consts(Q, &m, &l);
        const uint64_t m64 = ((((uint64_t)1 << (31u + l)) + Q - 1u) / Q);
        if (m64 > 0xFFFFFFFFull || (uint64_t)m != m64 || l > 31u) { ++bad; continue; }
        const uint32_t top = 0x7FFFFFFFu;
        const uint32_t kmax = top / Q;
        for (uint32_t j = 0; j < per_q + 16u; ++j) {
            uint64_t k;
            if (j < 8u) { k = kmax >= j ? kmax - j : 0u; }
            else if (j < 16u) { k = j - 8u; }
            else { s ^= s << 13; s ^= s >> 7; s ^= s << 17; k = s % ((uint64_t)kmax + 1u); }
            for (int d = -1; d <= 1; ++d) {
                const int64_t nn = (int64_t)(k * Q) + d;
                if (nn < 0 || nn > (int64_t)top) { continue; }
                const uint32_t n = (uint32_t)nn;
                if (tc_qdiv32(n, m, l) != n / Q) { ++bad; }
            }
        }
    }
    return bad;
}

/* exhaustive: every n < 2^31 for one divisor Q (returns mismatches) */
uint32_t mc_qdiv32_all(uint32_t Q)
{
    uint32_t m, l, bad = 0u;
    tc_qdiv32_consts(Q, &m, &l);
    for (uint32_t n = 0; n < 0x80000000u; ++n) { bad += tc_qdiv32(n, m, l) != n / Q; }
    return bad;
}

/* the decode kernel's packed context lookups against the plain bucket rules;
 * returns mismatches */
uint32_t mc_ctx_tables_check(void)
{
    uint32_t bad = 0u;
    for (uint32_t idx = 0; idx < 64u; ++idx) { bad += tc_ctx_pos_nib(idx) != tc_ctx_pos(idx); }
    for (uint32_t cat = 0; cat < 32u; ++cat) {
        const uint32_t w = cat > 1u ? cat - 1u : 0u, nz = cat != 0u;
        bad += tc_ctx_prev_wnz(w, nz) != tc_ctx_prev(cat, 1u);
        bad += 1u + tc_ctx_prev_wnz(w, nz) != tc_ctx_dc(cat, 1u);
    }
    return bad;
}

/* butterfly 8-point transforms against the direct products with the matrix
 * given in m[64] (row-major), for n pseudo-random vectors bounded by lim_f
 * (forward, int32) and lim_i (inverse); returns mismatches */
static const int mc_M[64] = {
    13, 13, 13, 13, 13, 13, 13, 13, 13, 11, 7, 1, -1, -7, -11, -13,
    12, 5, -5, -12, -12, -5, 5, 12, 11, -4, -13, -8, 8, 13, 4, -11,
    9, -9, -9, 9, 9, -9, -9, 9, 8, -13, 4, 11, -11, -4, 13, -8,
    5, -12, 12, -5, -5, 12, -12, 5, 1, -7, 11, -13, 13, -11, 7, -1,
};
uint32_t mc_dct8_check(const int32_t* m, uint32_t n, int32_t lim_f, int64_t lim_i)
{
    uint32_t bad = 0u;
    uint64_t s = 0x853C49E6748FEA9Bull;
    for (int i = 0; i < 64; ++i) { bad += m[i] != mc_M[i]; }
    for (uint32_t it = 0; it < n; ++it) {
        int32_t x[8], f[8];
        int64_t gi[8], oi[8];
        double gd[8], od[8];
        for (int k = 0; k < 8; ++k) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const int32_t e = it < 4u ? (it & 1u ? lim_f : -lim_f) : (int32_t)(s % (2u * (uint64_t)lim_f + 1u)) - lim_f;
            x[k] = (it == 2u || it == 3u) ? ((k & 1) ? -e : e) : e;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            const int64_t g = it < 4u ? (it & 1u ? lim_i : -lim_i) : (int64_t)(s % (2u * (uint64_t)lim_i + 1u)) - lim_i;
            gi[k] = (it == 2u || it == 3u) ? ((k & 1) ? -g : g) : g;
            gd[k] = (double)gi[k];
        }
        tc_fdct8_i32(x, f);
        tc_idct8_i64(gi, oi);
        tc_idct8_f64(gd, od);
        for (int v = 0; v < 8; ++v) {
            int64_t acc = 0, acci = 0;
            for (int k = 0; k < 8; ++k) { acc += (int64_t)mc_M[v * 8 + k] * x[k]; }
            for (int k = 0; k < 8; ++k) { acci += (int64_t)mc_M[k * 8 + v] * gi[k]; }
            bad += (int64_t)f[v] != acc;
            bad += oi[v] != acci;
            bad += od[v] != (double)acci;
        }
    }
    return bad;
}
