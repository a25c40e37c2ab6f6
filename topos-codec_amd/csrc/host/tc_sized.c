/* Target-size encode on the GPU (reference tc_frame_encode_sized,
 * native/include/topos_codec.h:315-328, native/src/codec/codec.c:7795-7890).
 *
 * The reference searches qp one probe at a time: each probe is an exact
 * count of the bytes a whole-frame encode at that qp writes (its DCT-once
 * F-cache path), and the search reads nothing but those byte counts. Here
 * the same decisions are replayed on the host over a table of byte counts,
 * and the GPU fills the table: whenever the replay needs a qp it has not
 * measured, one batched launch sequence encodes the frame at that qp and at
 * its unmeasured neighbours together (the entropy stage of one frame leaves
 * most of the device idle, so several candidate qps cost about one). The
 * replay restarts from the beginning after each launch; it is cheap, and it
 * keeps the budget, the monotonicity checks and the probe order exactly
 * the reference's. The packet of the chosen qp is one of the encoded
 * candidates. Searches of many frames (tcx_encode_sized_batch_dev) share
 * every launch.
 *
 * Search rules followed (codec.c):
 *   :7578-7787  strict (default): seeded start, log-slope jumps, bracket
 *               interpolation; the smallest qp whose packet fits
 *   :7199-7265  legacy linear (+8, +1, reclaim while below 3/4 of target)
 *   :7404-7551  r6 fast search (environment TOPOS_SIZED_SEARCH=r6)
 *   :7303-7351  24-probe budget, monotonicity violation -> legacy linear
 *   :6137-6143  exact probes need the 512 MiB F-cache; else legacy linear
 *   :7765-7785  seed state: the last two working points (qp, bytes) */
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "tc_sized.h"
#include "topos_gpu.h"

#define SZ_MAX_ITEMS 256u

/* ------------------------------------------------------------------ state */

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static tcx_sized_stream g_stream;          /* process-wide seed of tc_frame_encode_sized */
static atomic_int g_stream_busy;
static atomic_int g_mode = -1;             /* 0 strict, 1 linear, 2 fast; -1: read the environment */
static atomic_int g_seed_off, g_model_off;

void tcx_sized_reset(void)
{
    pthread_mutex_lock(&g_mu);
    memset(&g_stream, 0, sizeof(g_stream));
    pthread_mutex_unlock(&g_mu);
}

void tcx_sized_set_mode(int32_t mode) { atomic_store(&g_mode, mode < 0 || mode > 2 ? 0 : mode); }

void tcx_sized_set_seed(int32_t disable, int32_t model_disable)
{
    atomic_store(&g_seed_off, disable != 0);
    atomic_store(&g_model_off, model_disable != 0);
}

static int sz_mode(void)
{
    int m = atomic_load(&g_mode);
    if (m < 0) {
        const char* env = getenv("TOPOS_SIZED_SEARCH");
        m = env == NULL ? 0 : strcmp(env, "linear") == 0 ? 1 : strcmp(env, "r6") == 0 ? 2 : 0;
        int expect = -1;
        (void)atomic_compare_exchange_strong(&g_mode, &expect, m);
        m = atomic_load(&g_mode);
    }
    return m;
}

/* device memory for candidate packets and staged planes: grow-only chunks,
 * bump-allocated per call (one call at a time holds them) */
typedef struct sz_chunk { uint64_t base; size_t size, used; } sz_chunk;
static sz_chunk g_chunks[64];
static uint32_t g_nchunks;
static pthread_mutex_t g_pool_mu = PTHREAD_MUTEX_INITIALIZER;

static void pool_reset(void)
{
    for (uint32_t i = 0; i < g_nchunks; ++i) { g_chunks[i].used = 0; }
}

static int32_t pool_take(size_t bytes, uint64_t* out)
{
    bytes = (bytes + 1024u + 255u) & ~(size_t)255u;   /* 1 KiB tail slack for the kernels' wide loads */
    for (uint32_t i = 0; i < g_nchunks; ++i) {
        if (g_chunks[i].size - g_chunks[i].used >= bytes) {
            *out = g_chunks[i].base + g_chunks[i].used;
            g_chunks[i].used += bytes;
            return TC_OK;
        }
    }
    if (g_nchunks == 64u) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: device chunk table full"); }
    size_t sz = (size_t)256 << 20;
    while (sz < bytes) { sz *= 2u; }
    uint64_t p = 0;
    if (tcg_malloc(&p, sz) != 0) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: %s", tcg_last_error()); }
    g_chunks[g_nchunks] = (sz_chunk){ p, sz, bytes };
    g_nchunks++;
    *out = p;
    return TC_OK;
}

typedef struct sz_cand { uint32_t item, qp; } sz_cand;

/* encodes the candidates in one launch sequence and fills the tables; with
 * full = 1 every candidate gets a worst-case packet slot */
static int32_t sz_launch(tc_ectx* ec, sz_item* items, const sz_cand* cands, uint32_t n, int full)
{
    tc_enc_req* reqs = (tc_enc_req*)calloc(n, sizeof(tc_enc_req));
    topos_frame_config* cfgs = (topos_frame_config*)calloc(n, sizeof(topos_frame_config));
    if (reqs == NULL || cfgs == NULL) {
        free(reqs);
        free(cfgs);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: request table");
    }
    int32_t rc = TC_OK;
    for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
        sz_item* it = &items[cands[i].item];
        const uint32_t q = cands[i].qp;
        cfgs[i] = it->cfg;
        cfgs[i].qp_base = (uint8_t)q;
        const size_t cap = full ? it->full_cap : it->slot_cap;
        uint64_t slot = 0;
        rc = pool_take(cap, &slot);
        reqs[i].cfg = &cfgs[i];
        reqs[i].dev_planes = it->planes;
        reqs[i].dev_strides = it->strides;
        reqs[i].dev_out = slot;
        reqs[i].dev_out_cap = cap;
        reqs[i].stats = &it->st[q];
        it->pkt[q] = slot;
    }
    if (rc == TC_OK) {
        rc = tc_engine_encode(ec, reqs, n);
        if (rc == TC_ERR_BUFFER_TOO_SMALL || rc == TC_ERR_LIMIT_EXCEEDED || rc == TC_ERR_INVALID_ARGUMENT) {
            rc = TC_OK;   /* per-candidate outcomes below */
        }
    }
    for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
        sz_item* it = &items[cands[i].item];
        const uint32_t q = cands[i].qp;
        it->known[q] = 1;
        it->bytes[q] = reqs[i].packet_size;
        if (reqs[i].rc == TC_ERR_BUFFER_TOO_SMALL) {
            it->rc[q] = TC_OK;    /* measured; the packet did not fit the probe slot */
            it->stored[q] = 0;
        } else {
            it->rc[q] = reqs[i].rc;
            it->stored[q] = reqs[i].rc == TC_OK;
        }
    }
    free(reqs);
    free(cfgs);
    return rc;
}

/* runs the searches of n items to completion (caller holds the pool) */
static int32_t sz_run(tc_ectx* ec, sz_item* items, uint32_t n, int mode)
{
    sz_cand* cands = (sz_cand*)malloc(sizeof(sz_cand) * (size_t)n * 32u);
    if (cands == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: candidate list"); }
    int32_t rc = TC_OK;
    for (uint32_t round = 0; rc == TC_OK; ++round) {
        uint32_t missing = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (items[i].done) { continue; }
            const int32_t r = sz_replay(&items[i], mode);
            if (r == SZ_MISS) {
                missing++;
            } else {
                items[i].done = 1;
                items[i].result = r;
            }
        }
        if (missing == 0u) { break; }
        if (round > 4u * SZ_BUDGET) { rc = tc_fail(TC_ERR_STATE, "sized encode: search does not converge"); break; }
        /* speculation width: up to 9 qps for a lone search, fewer per frame
         * when many share a launch. The first round of a lone search without
         * a seed state takes 17: its second probe lands anywhere within the
         * log-slope jump, so a wide first round usually saves a launch (a
         * launch of 17 candidates of a 4K 422 HQ frame takes ~40 ms against
         * ~31 ms for 9; cold searches to 3 / 6 / 9 MB took 77 / 63 / 36 ms
         * with 9 and 61 / 42 / 43 ms with 17). Searches sharing a launch
         * stay narrow: 8 frames x 16 candidates made each round slower than
         * the launch it saved. */
        const uint32_t share = 64u / missing;
        uint32_t nc = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (items[i].done) { continue; }
            const int wide = round == 0u && missing == 1u && !(items[i].seed_on && items[i].seed.valid);
            uint32_t want = wide ? 17u : share;
            want = want < 3u ? 3u : want > (wide ? 17u : 9u) ? (wide ? 17u : 9u) : want;
            uint32_t qs[32];
            const uint32_t k = sz_candidates(&items[i], want, qs);
            for (uint32_t j = 0; j < k; ++j) { cands[nc++] = (sz_cand){ i, qs[j] }; }
        }
        rc = sz_launch(ec, items, cands, nc, 0);
    }
    /* finals whose packet did not fit a probe slot: encode them once more */
    if (rc == TC_OK) {
        uint32_t nc = 0;
        for (uint32_t i = 0; i < n; ++i) {
            sz_item* it = &items[i];
            const int32_t r = it->result;
            if ((r == TC_OK || r == TC_ERR_BUFFER_TOO_SMALL) && !it->stored[it->q_fin]) {
                cands[nc++] = (sz_cand){ i, it->q_fin };
            }
        }
        if (nc > 0u) { rc = sz_launch(ec, items, cands, nc, 1); }
    }
    free(cands);
    return rc;
}

static void sz_item_setup(sz_item* it, const topos_frame_config* cfg, uint32_t target, uint32_t qp_min,
                          uint32_t qp_max, size_t out_cap, const tcx_sized_stream* seed)
{
    tc_fhdr fh;
    tc_cfg_to_fhdr(cfg, &fh);
    (void)tc_fhdr_geometry(&fh);
    uint64_t blocks = 0;
    const uint32_t colour = fh.pixel_format == 3u ? 4u : 3u;
    for (uint32_t p = 0; p < colour && p < fh.plane_count; ++p) { blocks += (uint64_t)fh.bcols[p] * fh.brows[p]; }
    sz_item_init(it, cfg, target, qp_min, qp_max, out_cap, blocks > 0u && blocks * 64u * 4u <= SZ_FCACHE_MAX, seed,
                 !atomic_load(&g_seed_off), !atomic_load(&g_model_off));
    it->full_cap = tc_engine_encode_cap(cfg);
    const size_t want = (size_t)target + 4096u;
    it->slot_cap = want < it->full_cap ? want : it->full_cap;
}

/* sized-specific rules of the reference entry (codec.c:7795-7824) */
static int32_t sz_check(const topos_frame_config* cfg, uint32_t qp_min, uint32_t qp_max)
{
    int32_t rc = tc_cfg_validate(cfg);
    if (rc != TC_OK) { return rc; }
    const uint32_t em = cfg->reserved[0] & 0xFFu;
    if (em == 9u) { return tc_fail(TC_ERR_NOT_IMPLEMENTED, "segmented (V8) sized encode is not provided"); }
    if (em == 10u || em == 11u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "temporal sized encode is not a write path");
    }
    if (cfg->pixel_format == 3u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "RAW (pixel format 3) sized rate control is not provided");
    }
    if (qp_min > SZ_QP_MAX || qp_max > SZ_QP_MAX || qp_min > qp_max) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "qp range %u..%u invalid", qp_min, qp_max);
    }
    return TC_OK;
}

int32_t tc_frame_encode_sized(const topos_frame_config* cfg, const topos_frame_input* input, uint32_t target_bytes,
                              uint8_t qp_min, uint8_t qp_max, uint8_t* qp_used, uint8_t* out, size_t out_cap,
                              topos_frame_stats* stats)
{
    int32_t rc = sz_check(cfg, qp_min, qp_max);
    if (rc != TC_OK) { return rc; }
    if (out == NULL && out_cap != 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "out == NULL with cap %zu", out_cap);
    }
    if ((rc = tc_enc_check_host(cfg, input)) != TC_OK) {
        /* the legacy search reports its clamped start qp even on failure */
        if (qp_used != NULL) {
            const uint32_t q0 = cfg->qp_base < qp_min ? qp_min : cfg->qp_base > qp_max ? qp_max : cfg->qp_base;
            *qp_used = (uint8_t)(sz_mode() == 1 ? q0 : 0u);
        }
        return rc;
    }
    sz_item* it = (sz_item*)malloc(sizeof(sz_item));
    if (it == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: state"); }
    /* the seed state is process-wide; a concurrent caller searches without one */
    int expect = 0;
    const int own_seed = atomic_compare_exchange_strong(&g_stream_busy, &expect, 1);
    tcx_sized_stream seed;
    memset(&seed, 0, sizeof(seed));
    if (own_seed) {
        pthread_mutex_lock(&g_mu);
        seed = g_stream;
        pthread_mutex_unlock(&g_mu);
    }
    sz_item_setup(it, cfg, target_bytes, qp_min, qp_max, out_cap, &seed);
    tc_ectx* ec = NULL;
    if ((rc = tc_engine_acquire(&ec)) != TC_OK) {
        free(it);
        if (own_seed) { atomic_store(&g_stream_busy, 0); }
        return rc;
    }
    pthread_mutex_lock(&g_pool_mu);
    pool_reset();
    tc_fhdr fh;
    tc_cfg_to_fhdr(cfg, &fh);
    (void)tc_fhdr_geometry(&fh);
    for (uint32_t p = 0; p < fh.plane_count && rc == TC_OK; ++p) {
        const size_t row = (size_t)fh.pvw[p] * 2u;
        const size_t pitch = (((size_t)fh.pvw[p] + 7u) & ~(size_t)7u) * 2u;   /* 16-byte rows: vector loads */
        const size_t st = input->strides[p] ? input->strides[p] * 2u : row;
        rc = pool_take(pitch * fh.pvh[p], &it->planes[p]);
        if (rc == TC_OK && tcg_h2d_2d(tc_engine_stream(ec), it->planes[p], pitch, input->planes[p], st, row,
                                      fh.pvh[p]) != 0) {
            rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: plane upload: %s", tcg_last_error());
        }
        it->strides[p] = pitch / 2u;
    }
    if (rc == TC_OK) { rc = sz_run(ec, it, 1u, sz_mode()); }
    if (rc == TC_OK) {
        rc = it->result;
        const uint32_t q = it->q_fin;
        if ((rc == TC_OK || rc == TC_ERR_BUFFER_TOO_SMALL) && it->stored[q]) {
            if (stats != NULL) {
                *stats = it->st[q];
                stats->struct_size = (uint32_t)sizeof(topos_frame_stats);
                stats->abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
            }
            if (rc == TC_ERR_BUFFER_TOO_SMALL) {
                rc = tc_fail(TC_ERR_BUFFER_TOO_SMALL, "out_cap %zu < %u bytes needed", out_cap, it->bytes[q]);
            } else if (out != NULL && tcg_d2h_sync(tc_engine_stream(ec), out, it->pkt[q], it->bytes[q]) != 0) {
                rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "sized encode: packet download: %s", tcg_last_error());
            }
        }
        if (rc == TC_OK && it->seed_update && own_seed) {
            pthread_mutex_lock(&g_mu);
            g_stream = it->seed_out;
            pthread_mutex_unlock(&g_mu);
        }
        if (qp_used != NULL) { *qp_used = (uint8_t)it->q_used; }
    }
    pthread_mutex_unlock(&g_pool_mu);
    tc_engine_release(ec);
    if (own_seed) { atomic_store(&g_stream_busy, 0); }
    free(it);
    return rc;
}

int32_t tcx_encode_sized_batch_dev(tcx_sized_req* reqs, uint32_t n)
{
    if (reqs == NULL || n == 0u || n > SZ_MAX_ITEMS) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "sized batch: %u requests (1..%u)", n, SZ_MAX_ITEMS);
    }
    sz_item* items = (sz_item*)calloc(n, sizeof(sz_item));
    uint32_t* map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (items == NULL || map == NULL) {
        free(items);
        free(map);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "sized batch: state");
    }
    uint32_t m = 0;
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < n; ++i) {
        tcx_sized_req* r = &reqs[i];
        r->qp_used = 0;
        r->packet_size = 0;
        r->encodes = 0;
        int32_t rc = sz_check(&r->cfg, r->qp_min, r->qp_max);
        if (rc == TC_OK) { rc = tc_enc_check_dev(&r->cfg, r->dev_planes, r->dev_strides); }
        if (rc == TC_OK && r->dev_out == 0u) { rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "sized batch: dev_out == 0"); }
        r->rc = rc;
        if (rc != TC_OK) {
            if (first == TC_OK) { first = rc; }
            continue;
        }
        sz_item_setup(&items[m], &r->cfg, r->target_bytes, r->qp_min, r->qp_max, r->dev_out_cap, r->stream);
        memcpy(items[m].planes, r->dev_planes, sizeof(items[m].planes));
        tc_fhdr fh;
        tc_cfg_to_fhdr(&r->cfg, &fh);
        (void)tc_fhdr_geometry(&fh);
        for (uint32_t p = 0; p < 4u; ++p) {
            items[m].strides[p] = r->dev_strides[p] ? r->dev_strides[p] : fh.pvw[p];
        }
        map[m++] = i;
    }
    int32_t rc = TC_OK;
    if (m > 0u) {
        tc_ectx* ec = NULL;
        rc = tc_engine_acquire(&ec);
        if (rc == TC_OK) {
            pthread_mutex_lock(&g_pool_mu);
            pool_reset();
            rc = sz_run(ec, items, m, sz_mode());
            for (uint32_t k = 0; k < m && rc == TC_OK; ++k) {
                sz_item* it = &items[k];
                tcx_sized_req* r = &reqs[map[k]];
                const uint32_t q = it->q_fin;
                r->rc = it->result;
                r->qp_used = (uint8_t)it->q_used;
                r->encodes = it->count;
                if ((r->rc == TC_OK || r->rc == TC_ERR_BUFFER_TOO_SMALL) && it->stored[q]) {
                    r->packet_size = it->bytes[q];
                    r->stats = it->st[q];
                    r->stats.struct_size = (uint32_t)sizeof(topos_frame_stats);
                    r->stats.abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
                    if (r->rc == TC_OK && tcg_d2d(tc_engine_stream(ec), r->dev_out, it->pkt[q], it->bytes[q]) != 0) {
                        r->rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "sized batch: packet copy: %s", tcg_last_error());
                    }
                }
                if (r->rc == TC_OK && it->seed_update && r->stream != NULL) { *r->stream = it->seed_out; }
                if (first == TC_OK && r->rc != TC_OK) { first = r->rc; }
            }
            if (rc == TC_OK && tcg_sync(tc_engine_stream(ec)) != 0) {
                rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "sized batch: %s", tcg_last_error());
            }
            pthread_mutex_unlock(&g_pool_mu);
            tc_engine_release(ec);
        }
    }
    free(items);
    free(map);
    return rc != TC_OK ? rc : first;
}
