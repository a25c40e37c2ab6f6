/* Range allocator over the serving loop's HBM budget (tc_arena.c). */
#ifndef TC_ARENA_H
#define TC_ARENA_H

#include <stddef.h>
#include <stdint.h>

typedef struct tc_extent { size_t off, len; } tc_extent;

typedef struct tc_arena {
    size_t size;        /* bytes managed (multiple of 256) */
    size_t used, peak;  /* bytes held by allocations, and the most ever held */
    tc_extent* ext;     /* free extents sorted by offset, never adjacent */
    uint32_t n, cap;
} tc_arena;

int tc_arena_init(tc_arena* a, size_t size);
void tc_arena_fini(tc_arena* a);
/* the length an allocation of len bytes really holds */
size_t tc_arena_round(size_t len);
/* first fit; 0 on success, -1 when no free extent is large enough */
int tc_arena_alloc(tc_arena* a, size_t len, size_t* off);
/* returns a range (the len given to alloc); -1 on a range that is not held */
int tc_arena_free(tc_arena* a, size_t off, size_t len);
/* keeps the first keep bytes of a held range of len bytes, frees the rest */
int tc_arena_trim(tc_arena* a, size_t off, size_t len, size_t keep);
size_t tc_arena_largest(const tc_arena* a);

#endif
