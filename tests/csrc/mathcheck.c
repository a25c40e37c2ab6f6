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
