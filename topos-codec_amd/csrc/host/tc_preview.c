/* Target-size decodes: reduced (fixed ratio, low-frequency coefficients),
 * scaled (any target, all coefficients), the unified decode request with
 * AUTO_2K and DROP_ALPHA, and their batch / decoder-context forms (reference
 * native/include/topos_codec.h:341-380, 403-407, 441-497, 529-559;
 * native/src/codec/codec.c:9100-9640, 9640-9790). Argument rules, plan
 * resolution and output geometry on the host; the decode runs on the GPU: the
 * frame is reconstructed at source size keeping scan positions up to the
 * plan's coefficient limit, then sampled to the target (and deblocked on a
 * 4-sample grid for an exact 2:1 target of a deblocking stream). */
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

#define TC_BATCH_MAX_FRAMES 65536u
#define AUTO_2K_MAX_DIM 2048u

/* reference codec.c:9174-9194 */
int32_t tc_decode_scale_dimensions(uint32_t source_width, uint32_t source_height, tc_decode_scale scale,
                                   uint32_t* target_width, uint32_t* target_height)
{
    if (target_width == NULL || target_height == NULL || source_width == 0u || source_height == 0u ||
        (uint32_t)scale > (uint32_t)TC_DECODE_SCALE_EIGHTH) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "invalid decode scale dimensions");
    }
    static const uint32_t div[5] = { 1u, 2u, 3u, 4u, 8u };
    const uint32_t d = div[(uint32_t)scale];
    *target_width = (source_width + d - 1u) / d;
    *target_height = (source_height + d - 1u) / d;
    return TC_OK;
}

/* highest kept scan position per fixed ratio (codec.c:9196-9206) */
static uint32_t scale_limit(uint32_t scale)
{
    static const uint32_t lim[5] = { 63u, 24u, 16u, 10u, 4u };
    return scale <= 4u ? lim[scale] : 0u;
}

typedef struct dec_plan {
    uint32_t scaled, tw, th, limit, drop_alpha;
} dec_plan;

/* reference codec.c:9208-9273 */
int32_t tc_decode_request_validate(const topos_decode_request* q)
{
    if (q == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode request == NULL"); }
    if (q->struct_size != (uint32_t)sizeof(topos_decode_request) ||
        q->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode request struct_size/abi_version");
    }
    if (q->mode > TC_DECODE_MODE_AUTO_2K || q->quality > TC_DECODE_QUALITY_HIGH ||
        (q->memory_type != TC_DECODE_MEMORY_CPU && q->memory_type != TC_DECODE_MEMORY_HOST_VISIBLE_GPU) ||
        (q->flags & ~TC_DECODE_FLAG_DROP_ALPHA) != 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode request enum/flags invalid");
    }
    for (uint32_t i = 0; i < 4u; ++i) {
        if (q->reserved[i] != 0u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode request reserved nonzero"); }
    }
    if ((q->target_width == 0u) != (q->target_height == 0u)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode request target width/height must be paired");
    }
    switch (q->mode) {
    case TC_DECODE_MODE_FULL:
        if (q->scale != TC_DECODE_SCALE_FULL || q->target_width != 0u) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "full decode request has target/scale");
        }
        break;
    case TC_DECODE_MODE_SCALED:
        if (q->scale != TC_DECODE_SCALE_FULL || q->target_width == 0u) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "scaled decode request requires target");
        }
        break;
    case TC_DECODE_MODE_REDUCED:
        if (q->scale > TC_DECODE_SCALE_EIGHTH) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "reduced decode request scale invalid");
        }
        if (q->target_width != 0u) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "reduced decode request does not accept target");
        }
        break;
    default:
        if (q->scale != TC_DECODE_SCALE_FULL) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "auto2k decode request scale invalid");
        }
        break;
    }
    return TC_OK;
}

/* reference codec.c:9277-9357 */
static int32_t request_resolve(const tc_fhdr* fh, const topos_decode_request* q, dec_plan* plan)
{
    int32_t rc = tc_decode_request_validate(q);
    if (rc != TC_OK) { return rc; }
    memset(plan, 0, sizeof(*plan));
    plan->tw = fh->visible_w;
    plan->th = fh->visible_h;
    plan->limit = 63u;
    plan->drop_alpha = (q->flags & TC_DECODE_FLAG_DROP_ALPHA) != 0u && fh->alpha_mode != 0u;
    const uint32_t sw = fh->visible_w, sh = fh->visible_h;
    switch (q->mode) {
    case TC_DECODE_MODE_FULL:
        return TC_OK;
    case TC_DECODE_MODE_SCALED:
        plan->tw = q->target_width;
        plan->th = q->target_height;
        plan->scaled = 1u;
        break;
    case TC_DECODE_MODE_REDUCED:
        if ((rc = tc_decode_scale_dimensions(sw, sh, (tc_decode_scale)q->scale, &plan->tw, &plan->th)) != TC_OK) {
            return rc;
        }
        plan->scaled = q->scale != TC_DECODE_SCALE_FULL;
        plan->limit = scale_limit(q->scale);
        break;
    default:
        if (q->target_width != 0u) {
            plan->tw = q->target_width;
            plan->th = q->target_height;
            plan->scaled = plan->tw != sw || plan->th != sh;
            break;
        }
        if (sw <= AUTO_2K_MAX_DIM && sh <= AUTO_2K_MAX_DIM) { return TC_OK; }
        for (uint32_t s = TC_DECODE_SCALE_HALF; s <= TC_DECODE_SCALE_EIGHTH; ++s) {
            uint32_t tw = 0, th = 0;
            if ((rc = tc_decode_scale_dimensions(sw, sh, (tc_decode_scale)s, &tw, &th)) != TC_OK) { return rc; }
            if (tw <= AUTO_2K_MAX_DIM && th <= AUTO_2K_MAX_DIM) {
                plan->tw = tw;
                plan->th = th;
                plan->scaled = 1u;
                plan->limit = scale_limit(s);
                break;
            }
        }
        break;
    }
    if (plan->tw == 0u || plan->th == 0u || plan->tw > sw || plan->th > sh || plan->tw > 65535u ||
        plan->th > 65535u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode target %ux%u exceeds source %ux%u", plan->tw, plan->th, sw,
                       sh);
    }
    return TC_OK;
}

/* output geometry of a plan (codec.c:9359-9372) */
static void apply_plan_info(const dec_plan* plan, topos_frame_output* info)
{
    info->visible_width = (uint16_t)plan->tw;
    info->visible_height = (uint16_t)plan->th;
    info->coded_width = (uint16_t)plan->tw;
    info->coded_height = (uint16_t)plan->th;
    if (plan->drop_alpha && info->alpha_mode != 0u && info->plane_count >= 4u) {
        info->plane_count = 3u;
        info->alpha_mode = 0u;
        info->alpha_bit_depth = 0u;
    }
}

/* inter frames need a GOP context (codec.c:8868-8878) */
static int32_t p_guard(const tc_fhdr* fh)
{
    if (fh->major == 7u && fh->entropy_mode == 8u && fh->frame_type != 0u) {
        return tc_fail(TC_ERR_STATE, "inter frame decode requires a GOP context (stateless entry is I-only)");
    }
    return TC_OK;
}

/* output planes present and wide enough for the plan (codec.c:8581-8600) */
static int32_t check_outputs(const tc_view* v, const dec_plan* plan, uint32_t nplanes, uint16_t* const planes[4],
                             const size_t strides[4])
{
    for (uint32_t p = 0; p < nplanes; ++p) {
        if (planes[p] == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "planes_out[%u] == NULL", p); }
    }
    for (uint32_t p = 0; p < v->fh.plane_count; ++p) {
        if (plan->drop_alpha && tc_is_alpha(&v->fh, p)) { continue; }
        uint32_t w = v->fh.pvw[p], h = v->fh.pvh[p];
        if (plan->scaled) { tc_target_plane_dims(&v->fh, plan->tw, plan->th, p, &w, &h); }
        if (strides && strides[p] != 0u && strides[p] < w) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane %u stride %zu < visible width %u", p, strides[p], w);
        }
    }
    return TC_OK;
}

static void fill_req(tc_dec_req* r, const uint8_t* data, size_t size, const tc_view* v, const dec_plan* plan,
                     uint16_t* const* planes, const size_t* strides, topos_frame_output* info)
{
    memset(r, 0, sizeof(*r));
    r->data = data;
    r->size = size;
    r->view = v;
    r->planes = planes;
    r->strides = strides;
    r->info = info;
    if (plan->scaled) {
        r->tw = plan->tw;
        r->th = plan->th;
        r->limit = plan->limit;
    }
    r->skip_alpha = plan->drop_alpha;
}

static int32_t run_one(tc_dec_req* r)
{
    tc_ectx* ec = NULL;
    int32_t rc = tc_engine_acquire(&ec);
    if (rc != TC_OK) { return rc; }
    rc = tc_engine_decode(ec, r, 1u);
    tc_engine_release(ec);
    return rc >= TC_OK ? r->rc : rc;
}

/* one frame through a resolved plan; planes_out == NULL is query mode */
static int32_t decode_plan(const uint8_t* data, size_t size, tc_view* v, const dec_plan* plan,
                           uint16_t* const planes_out[4], const size_t strides[4], topos_frame_output* info)
{
    tc_fill_output_info(&v->fh, info);
    apply_plan_info(plan, info);
    if (planes_out == NULL) { return TC_OK; }
    int32_t rc = p_guard(&v->fh);
    if (rc == TC_OK) { rc = check_outputs(v, plan, info->plane_count, planes_out, strides); }
    if (rc != TC_OK) { return rc; }
    tc_dec_req r;
    fill_req(&r, data, size, v, plan, planes_out, strides, info);
    return run_one(&r);
}

static int32_t begin(topos_frame_output* info, const uint8_t* data)
{
    if (info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out_info == NULL"); }
    memset(info, 0, sizeof(*info));
    info->struct_size = (uint32_t)sizeof(topos_frame_output);
    info->abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
    if (data == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "data == NULL"); }
    return TC_OK;
}

int32_t tc_frame_decode_request(const uint8_t* data, size_t size, const topos_decode_request* request,
                                uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                                const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    int32_t rc = begin(out_info, data);
    if (rc == TC_OK) { rc = tc_decode_request_validate(request); }
    if (rc != TC_OK) { return rc; }
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    dec_plan plan;
    rc = tc_dec_prepare_check(data, size, out_info, v);
    if (rc == TC_OK) { rc = request_resolve(&v->fh, request, &plan); }
    if (rc == TC_OK) { rc = decode_plan(data, size, v, &plan, planes_out, strides, out_info); }
    free(v);
    return rc;
}

int32_t tc_frame_decode_scaled(const uint8_t* data, size_t size, uint32_t target_width, uint32_t target_height,
                               uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                               const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    int32_t rc = begin(out_info, data);
    if (rc != TC_OK) { return rc; }
    if (target_width == 0u || target_height == 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "scaled decode invalid arguments");
    }
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    rc = tc_dec_prepare_check(data, size, out_info, v);
    if (rc == TC_OK && (target_width > v->fh.visible_w || target_height > v->fh.visible_h)) {
        rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "scaled target %ux%u exceeds source %ux%u", target_width, target_height,
                     (unsigned)v->fh.visible_w, (unsigned)v->fh.visible_h);
    }
    if (rc == TC_OK) {
        const dec_plan plan = { 1u, target_width, target_height, 63u, 0u };
        rc = p_guard(&v->fh);
        if (rc == TC_OK) { rc = decode_plan(data, size, v, &plan, planes_out, strides, out_info); }
    }
    free(v);
    return rc;
}

int32_t tc_frame_decode_reduced(const uint8_t* data, size_t size, tc_decode_scale scale,
                                uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                                const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    if (out_info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out_info == NULL"); }
    if (scale == TC_DECODE_SCALE_FULL) { return tc_frame_decode(data, size, planes_out, strides, out_info); }
    if (data == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "data == NULL"); }
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    dec_plan plan = { 1u, 0u, 0u, scale_limit((uint32_t)scale), 0u };
    int32_t rc = begin(out_info, data);
    if (rc == TC_OK) { rc = tc_dec_prepare_check(data, size, out_info, v); }
    if (rc == TC_OK) {
        rc = tc_decode_scale_dimensions(v->fh.visible_w, v->fh.visible_h, scale, &plan.tw, &plan.th);
    }
    if (rc == TC_OK) {
        rc = p_guard(&v->fh);
        if (rc == TC_OK) { rc = decode_plan(data, size, v, &plan, planes_out, strides, out_info); }
    }
    free(v);
    return rc;
}

/* A batch through per-frame plans (codec.c:9640-9790); request NULL = the
 * fixed ratio `scale` for every frame (TC_DECODE_SCALE_FULL: the plain
 * tc_frame_decode_batch). The contract of topos_codec.h:374-400: a packet
 * whose structure does not parse rejects the whole batch before any frame is
 * decoded; an inter frame rejects the whole batch (TC_ERR_STATE); a frame
 * with a missing plane, a stride below its width or a bitstream generation
 * this library does not decode is rejected alone; every other frame is
 * decoded in the same launch sequence (concealed frames delivered) and the
 * first non-OK status in frame order is returned. */
int32_t tc_decode_batch_run(const topos_batch_packet* packets, uint32_t count, uint32_t scale,
                            const topos_decode_request* request, uint16_t* const* planes_out, const size_t* strides,
                            topos_frame_output* infos)
{
    if (count == 0u) { return TC_OK; }
    if (count > TC_BATCH_MAX_FRAMES) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "batch count %u > %u", count, TC_BATCH_MAX_FRAMES);
    }
    if (infos == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "infos == NULL"); }
    if (packets == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch packets == NULL"); }
    if (request == NULL && scale > (uint32_t)TC_DECODE_SCALE_EIGHTH) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "invalid decode scale %u", scale);
    }
    if (request != NULL) {
        int32_t rc = tc_decode_request_validate(request);
        if (rc != TC_OK) { return rc; }
    }
    for (uint32_t i = 0; i < count; ++i) {
        if (packets[i].data == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "packets[%u].data == NULL", i); }
    }
    tc_view* views = (tc_view*)malloc(sizeof(tc_view) * count);
    dec_plan* plans = (dec_plan*)calloc(count, sizeof(dec_plan));
    tc_dec_req* reqs = (tc_dec_req*)calloc(count, sizeof(tc_dec_req));
    uint32_t* map = (uint32_t*)calloc(count, sizeof(uint32_t));
    int32_t* rcs = (int32_t*)calloc(count, sizeof(int32_t));
    int32_t rc = TC_OK;
    if (views == NULL || plans == NULL || reqs == NULL || map == NULL || rcs == NULL) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "decode batch context");
    }
    /* structure of every packet first: one failure rejects the whole batch */
    for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
        rc = tc_dec_prepare(packets[i].data, packets[i].size, &infos[i], &views[i]);
        if (rc == TC_OK) {
            dec_plan* pl = &plans[i];
            if (request != NULL) {
                rc = request_resolve(&views[i].fh, request, pl);
            } else {
                memset(pl, 0, sizeof(*pl));
                pl->limit = scale_limit(scale);
                pl->scaled = scale != (uint32_t)TC_DECODE_SCALE_FULL;
                rc = tc_decode_scale_dimensions(views[i].fh.visible_w, views[i].fh.visible_h, (tc_decode_scale)scale,
                                                &pl->tw, &pl->th);
            }
        }
        if (rc == TC_OK) { apply_plan_info(&plans[i], &infos[i]); }
    }
    uint32_t nrun = 0;
    if (rc == TC_OK && planes_out != NULL) {
        for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
            rc = p_guard(&views[i].fh);
            if (rc != TC_OK) { break; }
            uint16_t* const* pl = planes_out + (size_t)i * TC_FRAME_MAX_PLANES;
            const size_t* st = strides ? strides + (size_t)i * TC_FRAME_MAX_PLANES : NULL;
            rcs[i] = tc_dec_supported(&views[i]);
            if (rcs[i] == TC_OK) { rcs[i] = check_outputs(&views[i], &plans[i], infos[i].plane_count, pl, st); }
            if (rcs[i] == TC_OK) {
                fill_req(&reqs[nrun], packets[i].data, packets[i].size, &views[i], &plans[i], pl, st, &infos[i]);
                map[nrun++] = i;
            }
        }
        if (rc == TC_OK && nrun > 0u) {
            tc_ectx* ec = NULL;
            int32_t erc = tc_engine_acquire(&ec);
            if (erc == TC_OK) {
                erc = tc_engine_decode(ec, reqs, nrun);
                tc_engine_release(ec);
            }
            for (uint32_t k = 0; k < nrun; ++k) { rcs[map[k]] = erc < TC_OK ? erc : reqs[k].rc; }
        }
        for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
            if (rcs[i] != TC_OK) { rc = rcs[i]; }
        }
    }
    free(views);
    free(plans);
    free(reqs);
    free(map);
    free(rcs);
    return rc;
}

int32_t tc_frame_decode_batch_reduced(const topos_batch_packet* packets, uint32_t count, tc_decode_scale scale,
                                      uint16_t* const* planes_out, const size_t* strides, topos_frame_output* infos)
{
    return tc_decode_batch_run(packets, count, (uint32_t)scale, NULL, planes_out, strides, infos);
}

/* device-resident batch through one request (no reference counterpart) */
int32_t tcx_decode_batch_dev_request(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                                     const uint64_t* dev_packets, const uint64_t* dev_planes,
                                     const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs,
                                     const topos_decode_request* request)
{
    if (n == 0u) { return TC_OK; }
    if (host_packets == NULL || sizes == NULL || dev_packets == NULL || dev_planes == NULL || infos == NULL ||
        rcs == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch args NULL");
    }
    int32_t rc = tc_decode_request_validate(request);
    if (rc != TC_OK) { return rc; }
    tc_view* views = (tc_view*)malloc(sizeof(tc_view) * n);
    dec_plan* plans = (dec_plan*)calloc(n, sizeof(dec_plan));
    tc_dec_req* reqs = (tc_dec_req*)calloc(n, sizeof(tc_dec_req));
    uint32_t* map = (uint32_t*)calloc(n, sizeof(uint32_t));
    if (views == NULL || plans == NULL || reqs == NULL || map == NULL) {
        free(views); free(plans); free(reqs); free(map);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch");
    }
    uint32_t nrun = 0;
    for (uint32_t i = 0; i < n; ++i) {
        rcs[i] = tc_dec_prepare_check(host_packets[i], sizes[i], &infos[i], &views[i]);
        if (rcs[i] == TC_OK) { rcs[i] = request_resolve(&views[i].fh, request, &plans[i]); }
        if (rcs[i] == TC_OK) { rcs[i] = p_guard(&views[i].fh); }
        if (rcs[i] != TC_OK) { continue; }
        apply_plan_info(&plans[i], &infos[i]);
        tc_dec_req* r = &reqs[nrun];
        fill_req(r, host_packets[i], sizes[i], &views[i], &plans[i], NULL, NULL, &infos[i]);
        r->dev_data = dev_packets[i];
        r->dev_planes = dev_planes + (size_t)i * 4u;
        r->dev_strides = dev_strides ? dev_strides + (size_t)i * 4u : NULL;
        map[nrun++] = i;
    }
    if (nrun > 0u) {
        tc_ectx* ec = NULL;
        int32_t erc = tc_engine_acquire(&ec);
        if (erc == TC_OK) {
            erc = tc_engine_decode(ec, reqs, nrun);
            tc_engine_release(ec);
        }
        for (uint32_t k = 0; k < nrun; ++k) { rcs[map[k]] = erc < TC_OK ? erc : reqs[k].rc; }
    }
    for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
        if (rcs[i] != TC_OK) { rc = rcs[i]; }
    }
    free(views); free(plans); free(reqs); free(map);
    return rc;
}
