/* CPU check of the target-size encode's search replay: the product's search
 * (topos-codec_amd/csrc/host/tc_sized_search.c, included verbatim) driven in
 * rounds like the GPU driver does, but with each candidate's byte count taken
 * from a table the test measured with the oracle encoder. Built by the tests
 * with gcc. */
#include <stdint.h>
#include <string.h>

#include "tc_sized_search.c"

/* res: rc, q_used, q_fin, size, count, rounds, measured, seed_update, seed_out[5] */
int sc_run(const uint32_t* bytes, const int32_t* rcs, const topos_frame_config* cfg, uint32_t target, uint32_t qmin,
           uint32_t qmax, uint64_t out_cap, int mode, int exact, const uint32_t* seed5, int seed_on, int model_on,
           uint32_t want, int32_t* res)
{
    static sz_item it;
    tcx_sized_stream seed;
    memset(&seed, 0, sizeof(seed));
    if (seed5 != NULL) {
        seed.qp = seed5[0]; seed.valid = seed5[1]; seed.bytes = seed5[2]; seed.prev = seed5[3];
        seed.prev_bytes = seed5[4];
    }
    sz_item_init(&it, cfg, target, qmin, qmax, (size_t)out_cap, exact, seed5 ? &seed : NULL, seed_on, model_on);
    uint32_t rounds = 0, measured = 0;
    int32_t r;
    for (;;) {
        r = sz_replay(&it, mode);
        if (r != SZ_MISS) { break; }
        if (rounds > 200u) { return -1; }
        uint32_t qs[SZ_QP_MAX + 1];
        const uint32_t k = sz_candidates(&it, want, qs);
        for (uint32_t j = 0; j < k; ++j) {
            if (qs[j] > SZ_QP_MAX || it.known[qs[j]]) { return -2; }
            it.known[qs[j]] = 1;
            it.bytes[qs[j]] = bytes[qs[j]];
            it.rc[qs[j]] = rcs[qs[j]];
            measured++;
        }
        rounds++;
    }
    res[0] = r; res[1] = (int32_t)it.q_used; res[2] = (int32_t)it.q_fin; res[3] = (int32_t)it.size;
    res[4] = (int32_t)it.count; res[5] = (int32_t)rounds; res[6] = (int32_t)measured; res[7] = it.seed_update;
    res[8] = (int32_t)it.seed_out.qp; res[9] = (int32_t)it.seed_out.valid; res[10] = (int32_t)it.seed_out.bytes;
    res[11] = (int32_t)it.seed_out.prev; res[12] = (int32_t)it.seed_out.prev_bytes;
    return 0;
}
