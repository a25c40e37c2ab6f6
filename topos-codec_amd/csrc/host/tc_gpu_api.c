/* Device-resident batch entry points (include/topos_gpu.h). */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

int32_t tcx_alpha_supported(void) { return 1; }

int32_t tcx_gpu_init(void)
{
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc == TC_OK) { tc_engine_release(c); }
    return rc;
}

int32_t tcx_dev_alloc(uint64_t* ptr, size_t bytes)
{
    if (ptr == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "ptr == NULL"); }
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc != TC_OK) { return rc; }
    rc = tcg_malloc(ptr, bytes + 1024u) == 0 ? TC_OK : tc_fail(TC_ERR_OUT_OF_MEMORY, "%s", tcg_last_error());
    tc_engine_release(c);
    return rc;
}

int32_t tcx_dev_free(uint64_t ptr)
{
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc != TC_OK) { return rc; }
    rc = tcg_free(ptr) == 0 ? TC_OK : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release(c);
    return rc;
}

int32_t tcx_h2d(uint64_t dst, const void* src, size_t bytes)
{
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc != TC_OK) { return rc; }
    rc = (tcg_h2d(tc_engine_stream(c), dst, src, bytes) == 0 && tcg_sync(tc_engine_stream(c)) == 0)
             ? TC_OK
             : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release(c);
    return rc;
}

int32_t tcx_d2h(void* dst, uint64_t src, size_t bytes)
{
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc != TC_OK) { return rc; }
    rc = tcg_d2h_sync(tc_engine_stream(c), dst, src, bytes) == 0
             ? TC_OK
             : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release(c);
    return rc;
}

int32_t tcx_copy(uint64_t dst, uint64_t src, size_t bytes)
{
    if (bytes % 16u != 0u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "copy size %zu is not a multiple of 16", bytes); }
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c);
    if (rc != TC_OK) { return rc; }
    rc = tcg_copy(tc_engine_stream(c), dst, src, bytes) == 0
             ? TC_OK
             : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release(c);
    return rc;
}

/* ---- batch preparation shared by the batch calls and the serve step ---- */

typedef struct enc_batch {
    uint32_t n, nrun;
    tc_enc_req* reqs;
    tc_enc_req* run;
    uint32_t* map;
} enc_batch;

typedef struct dec_batch {
    uint32_t n, nrun;
    tc_view* views;
    tc_dec_req* reqs;
    tc_dec_req* run;
    uint32_t* map;
} dec_batch;

static void enc_batch_free(enc_batch* b) { free(b->reqs); free(b->run); free(b->map); }
static void dec_batch_free(dec_batch* b) { free(b->views); free(b->reqs); free(b->run); free(b->map); }

static int32_t enc_batch_prepare(enc_batch* b, const topos_frame_config* cfgs, uint32_t n,
                                 const uint64_t* dev_planes, const size_t* dev_strides, const uint64_t* dev_out,
                                 const size_t* out_caps, topos_frame_stats* stats)
{
    memset(b, 0, sizeof(*b));
    if (n == 0u) { return TC_OK; }
    if (cfgs == NULL || dev_planes == NULL || dev_out == NULL || out_caps == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch encode args NULL");
    }
    extern int32_t tc_enc_check_dev(const topos_frame_config*, const uint64_t*, const size_t*);
    b->n = n;
    b->reqs = (tc_enc_req*)calloc(n, sizeof(tc_enc_req));
    b->run = (tc_enc_req*)calloc(n, sizeof(tc_enc_req));
    b->map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (b->reqs == NULL || b->run == NULL || b->map == NULL) {
        enc_batch_free(b);
        memset(b, 0, sizeof(*b));
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch");
    }
    for (uint32_t i = 0; i < n; ++i) {
        tc_enc_req* r = &b->reqs[i];
        r->cfg = &cfgs[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        r->dev_out = dev_out[i];
        r->dev_out_cap = out_caps[i];
        r->out_cap = out_caps[i];
        r->stats = stats ? &stats[i] : NULL;
        r->rc = tc_enc_check_dev(&cfgs[i], r->dev_planes, r->dev_strides);
        if (r->rc == TC_OK) { b->map[b->nrun] = i; b->run[b->nrun++] = *r; }
    }
    return TC_OK;
}

static int32_t enc_batch_report(enc_batch* b, int32_t rc, uint32_t* out_sizes, int32_t* rcs)
{
    for (uint32_t k = 0; k < b->nrun; ++k) {
        if (rc < TC_OK && b->run[k].rc == TC_OK) { b->run[k].rc = rc; }
        b->reqs[b->map[k]] = b->run[k];
    }
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < b->n; ++i) {
        if (rcs) { rcs[i] = b->reqs[i].rc; }
        if (out_sizes) { out_sizes[i] = b->reqs[i].packet_size; }
        if (first == TC_OK && b->reqs[i].rc != TC_OK) { first = b->reqs[i].rc; }
    }
    return first;
}

static int32_t dec_batch_prepare(dec_batch* b, uint32_t n, const uint8_t* const* host_packets,
                                 const size_t* sizes, const uint64_t* dev_packets, const uint64_t* dev_planes,
                                 const size_t* dev_strides, topos_frame_output* infos)
{
    memset(b, 0, sizeof(*b));
    if (n == 0u) { return TC_OK; }
    if (host_packets == NULL || sizes == NULL || dev_packets == NULL || dev_planes == NULL || infos == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch decode args NULL");
    }
    b->n = n;
    b->views = (tc_view*)malloc(sizeof(tc_view) * n);
    b->reqs = (tc_dec_req*)calloc(n, sizeof(tc_dec_req));
    b->run = (tc_dec_req*)calloc(n, sizeof(tc_dec_req));
    b->map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (b->views == NULL || b->reqs == NULL || b->run == NULL || b->map == NULL) {
        dec_batch_free(b);
        memset(b, 0, sizeof(*b));
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch");
    }
    for (uint32_t i = 0; i < n; ++i) {
        tc_dec_req* r = &b->reqs[i];
        r->rc = tc_dec_prepare_check(host_packets[i], sizes[i], &infos[i], &b->views[i]);
        r->data = host_packets[i];
        r->size = sizes[i];
        r->dev_data = dev_packets[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        r->info = &infos[i];
        r->view = &b->views[i];
        if (r->rc == TC_OK) { b->map[b->nrun] = i; b->run[b->nrun++] = *r; }
    }
    return TC_OK;
}

static int32_t dec_batch_report(dec_batch* b, int32_t rc, int32_t* rcs)
{
    for (uint32_t k = 0; k < b->nrun; ++k) { b->reqs[b->map[k]].rc = rc < TC_OK ? rc : b->run[k].rc; }
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < b->n; ++i) {
        if (rcs) { rcs[i] = b->reqs[i].rc; }
        if (first == TC_OK && b->reqs[i].rc != TC_OK) { first = b->reqs[i].rc; }
    }
    return first;
}

int32_t tcx_encode_batch_dev(const topos_frame_config* cfgs, uint32_t n, const uint64_t* dev_planes,
                             const size_t* dev_strides, const uint64_t* dev_out, const size_t* out_caps,
                             uint32_t* out_sizes, int32_t* rcs, topos_frame_stats* stats)
{
    enc_batch b;
    int32_t rc = enc_batch_prepare(&b, cfgs, n, dev_planes, dev_strides, dev_out, out_caps, stats);
    if (rc != TC_OK || n == 0u) { return rc; }
    if (b.nrun > 0u) {
        tc_ectx* c = NULL;
        rc = tc_engine_acquire(&c);
        if (rc == TC_OK) {
            rc = tc_engine_encode(c, b.run, b.nrun);
            tc_engine_release(c);
        }
    }
    const int32_t first = enc_batch_report(&b, rc, out_sizes, rcs);
    enc_batch_free(&b);
    return first;
}

int32_t tcx_decode_batch_dev(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                             const uint64_t* dev_packets, const uint64_t* dev_planes,
                             const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs)
{
    dec_batch b;
    int32_t rc = dec_batch_prepare(&b, n, host_packets, sizes, dev_packets, dev_planes, dev_strides, infos);
    if (rc != TC_OK || n == 0u) { return rc; }
    if (b.nrun > 0u) {
        tc_ectx* c = NULL;
        rc = tc_engine_acquire(&c);
        if (rc == TC_OK) {
            rc = tc_engine_decode(c, b.run, b.nrun);
            tc_engine_release(c);
        }
    }
    const int32_t first = dec_batch_report(&b, rc, rcs);
    dec_batch_free(&b);
    return first;
}

int32_t tcx_serve_step(const tcx_enc_batch* enc, const tcx_dec_batch* dec)
{
    enc_batch eb;
    dec_batch db;
    memset(&eb, 0, sizeof(eb));
    memset(&db, 0, sizeof(db));
    int32_t rc = TC_OK;
    if (enc != NULL) {
        rc = enc_batch_prepare(&eb, enc->cfgs, enc->n, enc->dev_planes, enc->dev_strides, enc->dev_out,
                               enc->out_caps, enc->stats);
    }
    if (rc == TC_OK && dec != NULL) {
        rc = dec_batch_prepare(&db, dec->n, dec->host_packets, dec->sizes, dec->dev_packets, dec->dev_planes,
                               dec->dev_strides, dec->infos);
    }
    if (rc != TC_OK) { enc_batch_free(&eb); dec_batch_free(&db); return rc; }
    tc_ectx* ce = NULL;
    tc_ectx* cd = NULL;
    int32_t rce = TC_OK, rcd = TC_OK;
    int enc_live = 0, dec_live = 0;
    if (eb.nrun > 0u) {
        rce = tc_engine_acquire(&ce);
        if (rce == TC_OK) { rce = tc_engine_encode_submit(ce, eb.run, eb.nrun); enc_live = rce == TC_OK; }
    }
    if (db.nrun > 0u) {
        rcd = tc_engine_acquire(&cd);
        if (rcd == TC_OK) { rcd = tc_engine_decode_submit(cd, db.run, db.nrun); dec_live = rcd == TC_OK; }
    }
    if (enc_live) { rce = tc_engine_encode_finish(ce, eb.run, eb.nrun); }
    if (dec_live) { rcd = tc_engine_decode_finish(cd, db.run, db.nrun); }
    if (ce != NULL) { tc_engine_release(ce); }
    if (cd != NULL) { tc_engine_release(cd); }
    int32_t first = TC_OK;
    if (enc != NULL) { first = enc_batch_report(&eb, rce, enc->out_sizes, enc->rcs); }
    if (dec != NULL) {
        const int32_t d = dec_batch_report(&db, rcd, dec->rcs);
        if (first == TC_OK) { first = d; }
    }
    enc_batch_free(&eb);
    dec_batch_free(&db);
    return first;
}

void tcx_set_timing(int on) { tcg_set_timing(on); }

void tcx_get_timing(float out[7])
{
    tcg_timing e, d;
    tc_engine_last_timing(&e, &d);
    out[0] = e.fdct_ms; out[1] = e.slice_enc_ms; out[2] = e.scan_ms; out[3] = e.emit_ms;
    out[4] = d.slice_dec_ms; out[5] = d.idct_ms; out[6] = d.deblock_ms;
}

int32_t tcx_device_info(uint32_t* cus, uint32_t* clock_khz)
{
    const int32_t rc = tcx_gpu_init();
    uint32_t c = 0u, k = 0u;
    if (rc == TC_OK) { tcg_device_info(&c, &k); }
    if (cus) { *cus = c; }
    if (clock_khz) { *clock_khz = k; }
    return rc;
}

void tcx_get_timing_pairs(uint64_t out[2])
{
    tcg_timing e, d;
    tc_engine_last_timing(&e, &d);
    out[0] = e.pairs;
    out[1] = d.pairs;
}

void tcx_set_graphs(int on) { tcg_set_graphs(on); }

void tcx_graph_stats(uint64_t* replays, uint64_t* captures)
{
    uint64_t h = 0, b = 0;
    tc_engine_graph_stats(&h, &b);
    if (replays) { *replays = h; }
    if (captures) { *captures = b; }
}

/* The no-allocation form of tcx_decode_batch_dev: the same launches, with
 * every intermediate (coefficients, descriptors, slice status) in the
 * library's static device scratch and every host array static, so a call
 * makes no device or host allocation (the first call in the process creates
 * its stream). Limits: TC_NA_MAX_FRAMES frames, 262,144 8x8 blocks (one
 * 3840x2160 4:2:2 frame) and TC_NA_MAX_SLICES slices per call, packets and
 * planes resident in HBM, full-resolution output; beyond them the call fails
 * with TC_ERR_LIMIT_EXCEEDED before any work. Calls are serialised by one lock:
 * the scratch holds the last call's contents. */
static pthread_mutex_t g_na_mu = PTHREAD_MUTEX_INITIALIZER;
static tc_view g_na_views[TC_NA_MAX_FRAMES];
static tc_dec_req g_na_reqs[TC_NA_MAX_FRAMES], g_na_run[TC_NA_MAX_FRAMES];
static uint32_t g_na_map[TC_NA_MAX_FRAMES];

int32_t tcx_decode_batch_dev_noalloc(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                                     const uint64_t* dev_packets, const uint64_t* dev_planes,
                                     const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs)
{
    if (n == 0u) { return TC_OK; }
    if (host_packets == NULL || sizes == NULL || dev_packets == NULL || dev_planes == NULL || infos == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch decode args NULL");
    }
    if (n > TC_NA_MAX_FRAMES) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "no-allocation decode: %u frames > %u", n, TC_NA_MAX_FRAMES);
    }
    pthread_mutex_lock(&g_na_mu);
    uint64_t blocks = 0u;
    uint32_t nrun = 0u, nslices = 0u;
    int32_t rc = TC_OK;
    for (uint32_t i = 0; i < n; ++i) {
        tc_dec_req* r = &g_na_reqs[i];
        memset(r, 0, sizeof(*r));
        r->rc = tc_dec_prepare_check(host_packets[i], sizes[i], &infos[i], &g_na_views[i]);
        r->data = host_packets[i];
        r->size = sizes[i];
        r->dev_data = dev_packets[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        r->info = &infos[i];
        r->view = &g_na_views[i];
        if (r->rc == TC_OK && r->dev_data == 0u) {
            r->rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "no-allocation decode: packet %u is not in device memory", i);
        }
        for (uint32_t p = 0; r->rc == TC_OK && p < g_na_views[i].fh.plane_count; ++p) {
            if (r->dev_planes[p] == 0u) {
                r->rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "no-allocation decode: frame %u plane %u is NULL", i, p);
            }
        }
        if (r->rc == TC_OK) {
            const tc_fhdr* fh = &g_na_views[i].fh;
            for (uint32_t p = 0; p < fh->plane_count; ++p) { blocks += (uint64_t)fh->bcols[p] * fh->brows[p]; }
            nslices += g_na_views[i].nslices;
            g_na_map[nrun] = i;
            g_na_run[nrun++] = *r;
        }
    }
    if (blocks * 256u > TCG_NOALLOC_Q_BYTES || nslices > TC_NA_MAX_SLICES) {
        rc = tc_fail(TC_ERR_LIMIT_EXCEEDED,
                     "no-allocation decode: %llu blocks / %u slices exceed %zu / %u", (unsigned long long)blocks,
                     nslices, TCG_NOALLOC_Q_BYTES / 256u, TC_NA_MAX_SLICES);
        pthread_mutex_unlock(&g_na_mu);
        return rc;
    }
    if (nrun > 0u) {
        tc_ectx* c = NULL;
        rc = tc_engine_acquire_noalloc(&c);
        if (rc == TC_OK) {
            rc = tc_engine_decode(c, g_na_run, nrun);
            tc_engine_release(c);
        }
    }
    for (uint32_t k = 0; k < nrun; ++k) { g_na_reqs[g_na_map[k]].rc = rc < TC_OK ? rc : g_na_run[k].rc; }
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < n; ++i) {
        if (rcs) { rcs[i] = g_na_reqs[i].rc; }
        if (first == TC_OK && g_na_reqs[i].rc != TC_OK) { first = g_na_reqs[i].rc; }
    }
    pthread_mutex_unlock(&g_na_mu);
    return first;
}
