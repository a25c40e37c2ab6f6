/* First-fit range allocator over one device allocation: the HBM budget of the
 * serving loop. Offsets and lengths are multiples of 256 bytes; the free list
 * is kept sorted by offset and coalesced on every free, so the arena never
 * holds two adjacent free extents. Not thread-safe: the caller serialises.
 */
#include <stdlib.h>
#include <string.h>

#include "tc_arena.h"

static size_t round256(size_t v) { return (v + 255u) & ~(size_t)255u; }

int tc_arena_init(tc_arena* a, size_t size)
{
    memset(a, 0, sizeof(*a));
    a->cap = 64u;
    a->ext = (tc_extent*)malloc(sizeof(tc_extent) * a->cap);
    if (a->ext == NULL) { return -1; }
    a->size = size & ~(size_t)255u;
    if (a->size > 0u) {
        a->ext[0].off = 0u;
        a->ext[0].len = a->size;
        a->n = 1u;
    }
    return 0;
}

void tc_arena_fini(tc_arena* a)
{
    free(a->ext);
    memset(a, 0, sizeof(*a));
}

size_t tc_arena_round(size_t len) { return round256(len ? len : 1u); }

int tc_arena_alloc(tc_arena* a, size_t len, size_t* off)
{
    len = round256(len ? len : 1u);
    for (uint32_t i = 0; i < a->n; ++i) {
        if (a->ext[i].len >= len) {
            *off = a->ext[i].off;
            a->ext[i].off += len;
            a->ext[i].len -= len;
            if (a->ext[i].len == 0u) {
                memmove(&a->ext[i], &a->ext[i + 1u], sizeof(tc_extent) * (a->n - i - 1u));
                a->n--;
            }
            a->used += len;
            if (a->used > a->peak) { a->peak = a->used; }
            return 0;
        }
    }
    return -1;
}

int tc_arena_free(tc_arena* a, size_t off, size_t len)
{
    len = round256(len ? len : 1u);
    if (off + len > a->size || (off & 255u) != 0u) { return -1; }
    /* first extent at or after off */
    uint32_t i = 0;
    while (i < a->n && a->ext[i].off < off) { ++i; }
    /* overlap with a free neighbour = double free */
    if (i > 0u && a->ext[i - 1u].off + a->ext[i - 1u].len > off) { return -1; }
    if (i < a->n && off + len > a->ext[i].off) { return -1; }
    const int left = i > 0u && a->ext[i - 1u].off + a->ext[i - 1u].len == off;
    const int right = i < a->n && off + len == a->ext[i].off;
    if (left && right) {
        a->ext[i - 1u].len += len + a->ext[i].len;
        memmove(&a->ext[i], &a->ext[i + 1u], sizeof(tc_extent) * (a->n - i - 1u));
        a->n--;
    } else if (left) {
        a->ext[i - 1u].len += len;
    } else if (right) {
        a->ext[i].off = off;
        a->ext[i].len += len;
    } else {
        if (a->n == a->cap) {
            tc_extent* e = (tc_extent*)realloc(a->ext, sizeof(tc_extent) * a->cap * 2u);
            if (e == NULL) { return -2; }
            a->ext = e;
            a->cap *= 2u;
        }
        memmove(&a->ext[i + 1u], &a->ext[i], sizeof(tc_extent) * (a->n - i));
        a->ext[i].off = off;
        a->ext[i].len = len;
        a->n++;
    }
    a->used -= len;
    return 0;
}

int tc_arena_trim(tc_arena* a, size_t off, size_t len, size_t keep)
{
    len = round256(len ? len : 1u);
    keep = round256(keep ? keep : 1u);
    if (keep >= len) { return 0; }
    return tc_arena_free(a, off + keep, len - keep);
}

size_t tc_arena_largest(const tc_arena* a)
{
    size_t m = 0u;
    for (uint32_t i = 0; i < a->n; ++i) { m = a->ext[i].len > m ? a->ext[i].len : m; }
    return m;
}
