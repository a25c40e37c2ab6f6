/* Target-size encode: the qp search replayed over measured byte counts
 * (internal to the host layer; tc_sized_search.c, driven by tc_sized.c). */
#ifndef TC_SIZED_H
#define TC_SIZED_H

#include <stddef.h>
#include <stdint.h>

#include "topos_codec.h"
#include "topos_gpu.h"

#define SZ_QP_MAX 95u
#define SZ_BUDGET 24u
#define SZ_FCACHE_MAX (512ull * 1024ull * 1024ull)
#define SZ_MISS 0x4D495353  /* the replay needs an unmeasured qp (it->miss) */

typedef struct sz_item {
    /* request */
    topos_frame_config cfg;
    uint64_t planes[4];
    size_t strides[4];
    uint32_t target, qp_min, qp_max;
    size_t out_cap;
    int exact;                  /* exact probes available (F-cache rule) */
    tcx_sized_stream seed;      /* seed state at the start */
    int seed_on, model_on;
    size_t slot_cap, full_cap;
    /* measured candidates */
    uint8_t known[SZ_QP_MAX + 1];
    uint8_t stored[SZ_QP_MAX + 1];  /* packet and stats kept in pkt[q] */
    uint32_t bytes[SZ_QP_MAX + 1];
    int32_t rc[SZ_QP_MAX + 1];
    uint64_t pkt[SZ_QP_MAX + 1];
    topos_frame_stats st[SZ_QP_MAX + 1];
    /* replay result */
    uint32_t miss;
    int done;
    int32_t result;
    uint32_t q_used;            /* qp reported to the caller */
    uint32_t q_fin;             /* qp of the final packet (also on BUFFER_TOO_SMALL) */
    uint32_t size, count;
    int seed_update;
    tcx_sized_stream seed_out;
} sz_item;

/* fills the request part of an item; exact = the frame's colour blocks fit
 * the 512 MiB F-cache (exact probes); seed NULL: no seed */
void sz_item_init(sz_item* it, const topos_frame_config* cfg, uint32_t target, uint32_t qp_min, uint32_t qp_max,
                  size_t out_cap, int exact, const tcx_sized_stream* seed, int seed_on, int model_on);
/* replays the whole search over the measured table (mode 0 strict, 1 legacy
 * linear, 2 r6); SZ_MISS or the search's status with q_used / q_fin / size /
 * count / seed_out filled */
int32_t sz_replay(sz_item* it, int mode);
/* the needed qp and up to want - 1 nearest unmeasured ones in the qp range */
uint32_t sz_candidates(const sz_item* it, uint32_t want, uint32_t* out);

#endif
