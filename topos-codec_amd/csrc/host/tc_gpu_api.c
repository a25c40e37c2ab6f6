/* Device-resident batch entry points (include/topos_gpu.h). */
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

int32_t tcx_alpha_supported(void) { return 1; }

int32_t tcx_gpu_init(void)
{
    int32_t rc = tc_engine_acquire();
    if (rc == TC_OK) { tc_engine_release(); }
    return rc;
}

int32_t tcx_dev_alloc(uint64_t* ptr, size_t bytes)
{
    if (ptr == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "ptr == NULL"); }
    int32_t rc = tc_engine_acquire();
    if (rc != TC_OK) { return rc; }
    rc = tcg_malloc(ptr, bytes + 1024u) == 0 ? TC_OK : tc_fail(TC_ERR_OUT_OF_MEMORY, "%s", tcg_last_error());
    tc_engine_release();
    return rc;
}

int32_t tcx_dev_free(uint64_t ptr)
{
    int32_t rc = tc_engine_acquire();
    if (rc != TC_OK) { return rc; }
    (void)tcg_sync();
    rc = tcg_free(ptr) == 0 ? TC_OK : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release();
    return rc;
}

int32_t tcx_h2d(uint64_t dst, const void* src, size_t bytes)
{
    int32_t rc = tc_engine_acquire();
    if (rc != TC_OK) { return rc; }
    rc = (tcg_h2d(dst, src, bytes) == 0 && tcg_sync() == 0) ? TC_OK
         : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release();
    return rc;
}

int32_t tcx_d2h(void* dst, uint64_t src, size_t bytes)
{
    int32_t rc = tc_engine_acquire();
    if (rc != TC_OK) { return rc; }
    rc = tcg_d2h(dst, src, bytes) == 0 ? TC_OK : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release();
    return rc;
}

int32_t tcx_copy(uint64_t dst, uint64_t src, size_t bytes)
{
    int32_t rc = tc_engine_acquire();
    if (rc != TC_OK) { return rc; }
    rc = tcg_copy(dst, src, bytes) == 0 ? TC_OK : tc_fail(TC_ERR_INVALID_ARGUMENT, "%s", tcg_last_error());
    tc_engine_release();
    return rc;
}

int32_t tcx_encode_batch_dev(const topos_frame_config* cfgs, uint32_t n, const uint64_t* dev_planes,
                             const size_t* dev_strides, const uint64_t* dev_out, const size_t* out_caps,
                             uint32_t* out_sizes, int32_t* rcs, topos_frame_stats* stats)
{
    if (n == 0u) { return TC_OK; }
    if (cfgs == NULL || dev_planes == NULL || dev_out == NULL || out_caps == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch encode args NULL");
    }
    extern int32_t tc_enc_check_dev(const topos_frame_config*, const uint64_t*, const size_t*);
    tc_enc_req* reqs = (tc_enc_req*)calloc(n, sizeof(tc_enc_req));
    tc_enc_req* run = (tc_enc_req*)calloc(n, sizeof(tc_enc_req));
    uint32_t* map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (reqs == NULL || run == NULL || map == NULL) {
        free(reqs); free(run); free(map);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch");
    }
    uint32_t nrun = 0;
    for (uint32_t i = 0; i < n; ++i) {
        tc_enc_req* r = &reqs[i];
        r->cfg = &cfgs[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        r->dev_out = dev_out[i];
        r->dev_out_cap = out_caps[i];
        r->out_cap = out_caps[i];
        r->stats = stats ? &stats[i] : NULL;
        r->rc = tc_enc_check_dev(&cfgs[i], r->dev_planes, r->dev_strides);
        if (r->rc == TC_OK) { map[nrun] = i; run[nrun++] = *r; }
    }
    int32_t rc = TC_OK;
    if (nrun > 0u) {
        rc = tc_engine_acquire();
        if (rc == TC_OK) {
            rc = tc_engine_encode(run, nrun);
            tc_engine_release();
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]] = run[k]; }
        } else {
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]].rc = rc; }
        }
    }
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < n; ++i) {
        if (rcs) { rcs[i] = reqs[i].rc; }
        if (out_sizes) { out_sizes[i] = reqs[i].packet_size; }
        if (first == TC_OK && reqs[i].rc != TC_OK) { first = reqs[i].rc; }
    }
    free(reqs); free(run); free(map);
    return first;
}

int32_t tcx_decode_batch_dev(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                             const uint64_t* dev_packets, const uint64_t* dev_planes,
                             const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs)
{
    if (n == 0u) { return TC_OK; }
    if (host_packets == NULL || sizes == NULL || dev_packets == NULL || dev_planes == NULL || infos == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch decode args NULL");
    }
    tc_view* views = (tc_view*)malloc(sizeof(tc_view) * n);
    tc_dec_req* reqs = (tc_dec_req*)calloc(n, sizeof(tc_dec_req));
    tc_dec_req* run = (tc_dec_req*)calloc(n, sizeof(tc_dec_req));
    uint32_t* map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (views == NULL || reqs == NULL || run == NULL || map == NULL) {
        free(views); free(reqs); free(run); free(map);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch");
    }
    extern int32_t tc_dec_prepare_check(const uint8_t*, size_t, topos_frame_output*, tc_view*);
    uint32_t nrun = 0;
    for (uint32_t i = 0; i < n; ++i) {
        tc_dec_req* r = &reqs[i];
        r->rc = tc_dec_prepare_check(host_packets[i], sizes[i], &infos[i], &views[i]);
        r->data = host_packets[i];
        r->size = sizes[i];
        r->dev_data = dev_packets[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        r->info = &infos[i];
        r->view = &views[i];
        if (r->rc == TC_OK) { map[nrun] = i; run[nrun++] = *r; }
    }
    int32_t rc = TC_OK;
    if (nrun > 0u) {
        rc = tc_engine_acquire();
        if (rc == TC_OK) {
            rc = tc_engine_decode(run, nrun);
            tc_engine_release();
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]].rc = rc < TC_OK ? rc : run[k].rc; }
        } else {
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]].rc = rc; }
        }
    }
    int32_t first = TC_OK;
    for (uint32_t i = 0; i < n; ++i) {
        if (rcs) { rcs[i] = reqs[i].rc; }
        if (first == TC_OK && reqs[i].rc != TC_OK) { first = reqs[i].rc; }
    }
    free(views); free(reqs); free(run); free(map);
    return first;
}

void tcx_set_timing(int on) { tcg_set_timing(on); }

void tcx_get_timing(float out[7])
{
    tcg_timing t;
    tcg_get_timing(&t);
    out[0] = t.fdct_ms; out[1] = t.slice_enc_ms; out[2] = t.scan_ms; out[3] = t.emit_ms;
    out[4] = t.slice_dec_ms; out[5] = t.idct_ms; out[6] = t.deblock_ms;
}
