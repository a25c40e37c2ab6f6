/* CPU check of the serving loop's HBM range allocator: random alloc / free /
 * trim sequences against a byte-ownership model. Built by the tests with gcc
 * straight from the product source (topos-codec_amd/csrc/host/tc_arena.c). */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "tc_arena.c"

static uint64_t rng_state;
static uint64_t rnd(void)
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* returns 0 when every invariant held; otherwise the failing step + 1 */
int ac_run(uint64_t seed, uint32_t steps, uint32_t units)
{
    rng_state = seed;
    tc_arena a;
    const size_t size = (size_t)units * 256u;
    if (tc_arena_init(&a, size) != 0) { return -1; }
    int32_t* owner = (int32_t*)malloc(sizeof(int32_t) * units);   /* -1 free, else allocation id */
    size_t* offs = (size_t*)malloc(sizeof(size_t) * steps);
    size_t* lens = (size_t*)malloc(sizeof(size_t) * steps);
    uint32_t* live = (uint32_t*)malloc(sizeof(uint32_t) * steps);
    uint32_t nlive = 0;
    for (uint32_t i = 0; i < units; ++i) { owner[i] = -1; }
    int bad = 0;
    for (uint32_t s = 0; s < steps && !bad; ++s) {
        const uint32_t op = (uint32_t)(rnd() % 8u);
        if (op < 4u || nlive == 0u) {
            const size_t len = 1u + (size_t)(rnd() % (size / 6u));
            size_t off = 0;
            /* model: is there a free run of round(len) units? (first fit) */
            const uint32_t need = (uint32_t)(tc_arena_round(len) / 256u);
            uint32_t run = 0, first = UINT32_MAX;
            for (uint32_t u = 0; u < units; ++u) {
                run = owner[u] < 0 ? run + 1u : 0u;
                if (run == need) { first = u + 1u - need; break; }
            }
            const int rc = tc_arena_alloc(&a, len, &off);
            if ((rc == 0) != (first != UINT32_MAX)) { bad = (int)s + 1; break; }
            if (rc == 0) {
                if (off != (size_t)first * 256u) { bad = (int)s + 1; break; }
                for (uint32_t u = 0; u < need; ++u) { owner[first + u] = (int32_t)s; }
                offs[s] = off; lens[s] = len; live[nlive++] = s;
            }
        } else if (op < 7u) {
            const uint32_t k = (uint32_t)(rnd() % nlive);
            const uint32_t id = live[k];
            if (tc_arena_free(&a, offs[id], lens[id]) != 0) { bad = (int)s + 1; break; }
            /* a second free of the same range must be refused */
            if (tc_arena_free(&a, offs[id], lens[id]) == 0) { bad = (int)s + 1; break; }
            for (uint32_t u = 0; u < units; ++u) { if (owner[u] == (int32_t)id) { owner[u] = -1; } }
            live[k] = live[--nlive];
        } else {
            const uint32_t k = (uint32_t)(rnd() % nlive);
            const uint32_t id = live[k];
            const size_t keep = 1u + (size_t)(rnd() % lens[id]);
            if (tc_arena_trim(&a, offs[id], lens[id], keep) != 0) { bad = (int)s + 1; break; }
            const uint32_t kunits = (uint32_t)(tc_arena_round(keep) / 256u);
            uint32_t seen = 0;
            for (uint32_t u = 0; u < units; ++u) {
                if (owner[u] == (int32_t)id) { if (seen++ >= kunits) { owner[u] = -1; } }
            }
            lens[id] = keep;
        }
        /* invariants: used bytes, sorted non-adjacent free extents covering exactly the free units */
        size_t used = 0;
        for (uint32_t u = 0; u < units; ++u) { used += owner[u] >= 0 ? 256u : 0u; }
        if (used != a.used || a.peak < a.used) { bad = (int)s + 1; break; }
        size_t freeb = 0;
        for (uint32_t i = 0; i < a.n; ++i) {
            if (i > 0u && a.ext[i - 1u].off + a.ext[i - 1u].len >= a.ext[i].off) { bad = (int)s + 1; break; }
            for (size_t o = a.ext[i].off; o < a.ext[i].off + a.ext[i].len; o += 256u) {
                if (owner[o / 256u] >= 0) { bad = (int)s + 1; break; }
            }
            freeb += a.ext[i].len;
        }
        if (!bad && freeb + used != size) { bad = (int)s + 1; }
    }
    free(owner); free(offs); free(lens); free(live);
    tc_arena_fini(&a);
    return bad;
}
