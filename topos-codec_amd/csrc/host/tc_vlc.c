/* V2 canonical VLC books (codebook version 1), built once on the host and
 * uploaded to the device in the lane layouts the slice kernels use.
 *
 * The code lengths are frozen bitstream data
 * (native/src/entropy/vlc_tables.h:28-67): DC category 29 symbols, run 64
 * (63 = end of block), level category 28 with category 0 outside the
 * alphabet. Codes are canonical: lengths ascending, symbols ascending within
 * one length, consecutive code values (native/src/entropy/vlc.c:37-53).
 */
#include <string.h>

#include "tc_host.h"

static const uint8_t k_dc_len[4][29] = {
    { 3, 13, 13, 13, 13, 13, 13, 12, 8, 6, 6, 5, 4, 3, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 7, 7, 7, 9, 10 },
    { 3, 10, 15, 15, 14, 14, 11, 7, 6, 6, 5, 4, 3, 3, 3, 3, 4, 4, 5, 5, 6, 7, 7, 7, 7, 8, 9, 12, 14 },
    { 3, 10, 13, 11, 8, 7, 6, 5, 4, 3, 3, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 7, 7, 9, 12, 15, 15, 15, 15 },
    { 3, 6, 5, 4, 4, 3, 3, 3, 4, 4, 4, 5, 5, 5, 6, 7, 7, 7, 8, 9, 10, 11, 14, 14, 14, 14, 14, 14, 13 },
};

static const uint8_t k_run_len[4][64] = {
    { 1, 2, 5, 6, 7, 6, 7, 7, 8, 7, 7, 8, 9, 7, 8, 13, 15, 10, 16, 20, 20, 12, 12, 14, 13, 20, 17, 20, 15, 20, 20,
      20, 14, 20, 20, 20, 20, 20, 18, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 20, 20, 3 },
    { 1, 2, 5, 6, 7, 6, 7, 8, 7, 7, 7, 8, 9, 7, 8, 13, 15, 10, 16, 20, 20, 12, 12, 13, 13, 20, 17, 20, 16, 20, 20,
      20, 16, 20, 20, 20, 20, 20, 18, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 20, 20, 3 },
    { 1, 2, 4, 5, 7, 7, 8, 9, 10, 9, 9, 10, 11, 9, 10, 15, 16, 12, 18, 19, 19, 14, 14, 16, 16, 20, 20, 20, 19, 20,
      20, 20, 19, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 20, 20, 20, 3 },
    { 1, 2, 4, 5, 5, 6, 6, 7, 8, 8, 8, 9, 9, 9, 9, 11, 11, 10, 11, 12, 12, 12, 13, 13, 13, 13, 14, 15, 15, 16, 16,
      15, 16, 17, 17, 18, 19, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 20, 20, 4 },
};

static const uint8_t k_lvl_len[4][28] = {
    { 0, 10, 9, 8, 7, 5, 4, 4, 3, 3, 3, 3, 3, 4, 5, 5, 6, 6, 7, 8, 8, 9, 8, 7, 7, 7, 9, 10 },
    { 0, 6, 5, 4, 4, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 8, 8, 7, 7, 7, 9, 10, 11, 12, 12 },
    { 0, 4, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 6, 6, 7, 8, 7, 7, 6, 7, 7, 9, 10, 13, 13, 12, 12, 12 },
    { 0, 3, 2, 3, 3, 3, 4, 5, 5, 6, 6, 6, 6, 6, 6, 6, 7, 8, 9, 10, 13, 13, 13, 13, 13, 13, 13, 13 },
};

const uint8_t* tc_vlc_lengths(uint32_t family, uint32_t book, uint32_t* nsym)
{
    if (book > 3u || family > 2u) { return NULL; }
    if (nsym) { *nsym = family == 0u ? 29u : family == 1u ? 64u : 28u; }
    return family == 0u ? k_dc_len[book] : family == 1u ? k_run_len[book] : k_lvl_len[book];
}

/* One book: codes by canonical assignment; for the decoder, per length l the
 * left-aligned (20-bit) end of the length-l code range, so that the length
 * of the next code is the smallest l whose end exceeds the next 20 bits, and
 * the first code / first order index of that length. */
static int book_build(tcg_vlc_books* b, uint32_t slot, const uint8_t* len, uint32_t nsym)
{
    uint32_t code = 0, n = 0;
    uint64_t kraft = 0;
    for (uint32_t s = 0; s < nsym; ++s) {
        if (len[s] > 20u) { return -1; }
        if (len[s]) { kraft += 1ull << (20u - len[s]); }
    }
    if (kraft != (1ull << 20)) { return -1; } /* the frozen books are complete prefix codes */
    for (uint32_t l = 1; l <= 20u; ++l) {
        const uint32_t first = code, idx = n;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (len[s] != l) { continue; }
            b->code[slot][s] = code++;
            b->len[slot][s] = l;
            b->sym[slot][n++] = s;
        }
        b->info[slot][l] = (first << 8) | idx;
        b->lim[slot][l] = code << (20u - l);
        code <<= 1;
    }
    return 0;
}

int32_t tc_vlc_books_build(tcg_vlc_books* b)
{
    memset(b, 0, sizeof(*b));
    for (uint32_t f = 0; f < 3u; ++f) {
        for (uint32_t k = 0; k < 4u; ++k) {
            uint32_t nsym = 0;
            const uint8_t* len = tc_vlc_lengths(f, k, &nsym);
            if (book_build(b, f * 4u + k, len, nsym) != 0) {
                return tc_fail(TC_ERR_STATE, "VLC book %u/%u is not a complete prefix code", f, k);
            }
        }
    }
    return TC_OK;
}
