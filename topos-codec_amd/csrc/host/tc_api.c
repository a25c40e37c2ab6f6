/* Public frame entry points of include/topos_codec.h on top of the GPU engine.
 * Argument checks and their order follow the reference entry points
 * (native/src/codec/codec.c:5730-5808 encode setup, 5970-6018 tc_frame_encode,
 * 9005-9098 tc_frame_decode, 9639-9805 batch decode, 9807-9833 geometry;
 * native/src/codec/decoder_ctx.c for the context calls).
 */
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

static uint32_t pad8(uint32_t v) { return (v + 7u) / 8u * 8u; }

int32_t tc_frame_config_validate(const topos_frame_config* cfg)
{
    return tc_cfg_validate(cfg);
}

/* allocation hint: 1024 B per colour block, 16 B per alpha pixel, capped at
 * 256 MiB (native/src/codec/codec.c:607-650) */
size_t tc_frame_packet_bound(const topos_frame_config* cfg)
{
    if (cfg == NULL || cfg->visible_width == 0u || cfg->visible_height == 0u ||
        cfg->visible_width > TC_MAX_DIM || cfg->visible_height > TC_MAX_DIM) {
        return 0u;
    }
    const uint64_t cw = pad8(cfg->visible_width), ch = pad8(cfg->visible_height);
    uint64_t cwc, chc;
    if (cfg->pixel_format == 0u) {
        cwc = pad8((uint32_t)(((uint64_t)cfg->visible_width + 1u) / 2u)); chc = ch;
    } else if (cfg->pixel_format == 3u) {
        cwc = pad8((uint32_t)(((uint64_t)cfg->visible_width + 1u) / 2u));
        chc = pad8((uint32_t)(((uint64_t)cfg->visible_height + 1u) / 2u));
    } else {
        cwc = cw; chc = ch;
    }
    const uint64_t blocks = cfg->pixel_format == 3u ? 4u * ((cwc * chc) / 64u) : (cw * ch) / 64u + 2u * ((cwc * chc) / 64u);
    const uint64_t planes = cfg->pixel_format == 3u ? 4u : 3u;
    uint64_t bound = 53u + planes * (17u + blocks * 1024u);
    if (cfg->alpha_mode != 0u) { bound += 17u + cw * ch * 16u; }
    return bound > TC_MAX_PACKET ? (size_t)TC_MAX_PACKET : (size_t)bound;
}

/* encode-side argument checks shared by the host and the resident entries */
static int32_t enc_check(const topos_frame_config* cfg, const topos_frame_input* input,
                         const uint64_t* dev_planes, const size_t* dev_strides)
{
    int32_t rc = tc_cfg_validate(cfg);
    if (rc != TC_OK) { return rc; }
    tc_fhdr fh;
    tc_cfg_to_fhdr(cfg, &fh);
    (void)tc_fhdr_geometry(&fh);
    if (dev_planes == NULL) {
        if (input == NULL || (input->struct_size != 0u && input->struct_size != (uint32_t)sizeof(topos_frame_input))) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "input struct invalid");
        }
    }
    for (uint32_t p = 0; p < fh.plane_count; ++p) {
        const void* pl = dev_planes ? (const void*)(uintptr_t)dev_planes[p] : (const void*)input->planes[p];
        const size_t st = dev_planes ? (dev_strides ? dev_strides[p] : 0u) : input->strides[p];
        if (pl == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "input plane %u == NULL", p); }
        if (st != 0u && st < fh.pvw[p]) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "input stride[%u]=%zu < visible width %u", p, st,
                           (unsigned)fh.pvw[p]);
        }
    }
    const uint32_t em = cfg->reserved[0] & 0xFFu;
    if (cfg->reserved[1] == 1u || cfg->reserved[2] == 1u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "adaptive quantisation / RDO encode is not provided by the GPU backend");
    }
    if (em != 8u && em != 0u && em != 1u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED,
                       "entropy mode %u is not provided by the GPU backend (V1 Rice = 0, V2 VLC = 1, rANS2 = 8)",
                       (unsigned)em);
    }
    if (cfg->alpha_mode != 0u && !tcx_alpha_supported()) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "alpha planes are not provided by this build");
    }
    return TC_OK;
}

int32_t tc_enc_check_dev(const topos_frame_config* cfg, const uint64_t* dev_planes, const size_t* dev_strides)
{
    return enc_check(cfg, NULL, dev_planes, dev_strides);
}

int32_t tc_enc_check_host(const topos_frame_config* cfg, const topos_frame_input* input)
{
    return enc_check(cfg, input, NULL, NULL);
}

int32_t tc_frame_encode(const topos_frame_config* cfg, const topos_frame_input* input,
                        uint8_t* out, size_t out_cap, topos_frame_stats* stats)
{
    topos_frame_stats local;
    memset(&local, 0, sizeof(local));
    if (stats == NULL) { stats = &local; }
    stats->struct_size = (uint32_t)sizeof(topos_frame_stats);
    stats->abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
    if (out == NULL && out_cap != 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "out == NULL with cap %zu", out_cap);
    }
    int32_t rc = enc_check(cfg, input, NULL, NULL);
    if (rc != TC_OK) { return rc; }
    tc_ectx* ec = NULL;
    if ((rc = tc_engine_acquire(&ec)) != TC_OK) { return rc; }
    tc_enc_req r;
    memset(&r, 0, sizeof(r));
    r.cfg = cfg;
    r.input = input;
    r.out = out;
    r.out_cap = out_cap;
    r.stats = stats;
    rc = tc_engine_encode(ec, &r, 1u);
    tc_engine_release(ec);
    return rc == TC_OK ? r.rc : rc;
}

static int32_t dec_prepare(const uint8_t* data, size_t size, topos_frame_output* info, tc_view* v)
{
    if (info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out_info == NULL"); }
    memset(info, 0, sizeof(*info));
    info->struct_size = (uint32_t)sizeof(topos_frame_output);
    info->abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
    if (data == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "data == NULL"); }
    int32_t rc = tc_parse_structure(data, size, v);
    if (rc != TC_OK) { return rc; }
    tc_fill_output_info(&v->fh, info);
    return TC_OK;
}

static int32_t dec_supported(const tc_view* v)
{
    const tc_fhdr* fh = &v->fh;
    if (fh->major == 7u && fh->entropy_mode == 8u && fh->frame_type != 0u) {
        return tc_fail(TC_ERR_STATE, "inter frame decode requires a GOP context");
    }
    if (!(fh->major == 7u && fh->entropy_mode == 7u) && fh->major != 1u &&
        !(fh->major == 2u && fh->entropy_mode == 1u)) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED,
                       "bitstream generation (%u, entropy %u) is not decoded by the GPU backend",
                       (unsigned)fh->major, (unsigned)fh->entropy_mode);
    }
    if (fh->alpha_mode != 0u && !tcx_alpha_supported()) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "alpha planes are not provided by this build");
    }
    return TC_OK;
}

int32_t tc_dec_prepare_check(const uint8_t* data, size_t size, topos_frame_output* info, tc_view* v)
{
    int32_t rc = dec_prepare(data, size, info, v);
    return rc == TC_OK ? dec_supported(v) : rc;
}

int32_t tc_frame_decode(const uint8_t* data, size_t size, uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                        const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    int32_t rc = dec_prepare(data, size, out_info, v);
    if (rc == TC_OK && planes_out != NULL) {
        rc = dec_supported(v);
        for (uint32_t p = 0; rc == TC_OK && p < v->fh.plane_count; ++p) {
            if (planes_out[p] == NULL) { rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "planes_out[%u] == NULL", p); }
            else if (strides && strides[p] != 0u && strides[p] < v->fh.pvw[p]) {
                rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "plane %u stride %zu < visible width %u", p, strides[p],
                             (unsigned)v->fh.pvw[p]);
            }
        }
        tc_ectx* ec = NULL;
        if (rc == TC_OK && (rc = tc_engine_acquire(&ec)) == TC_OK) {
            tc_dec_req r;
            memset(&r, 0, sizeof(r));
            r.data = data; r.size = size; r.planes = planes_out; r.strides = strides;
            r.info = out_info; r.view = v;
            rc = tc_engine_decode(ec, &r, 1u);
            tc_engine_release(ec);
            if (rc >= TC_OK) { rc = r.rc; }
        }
    }
    free(v);
    return rc;
}

int32_t tc_frame_decode_batch(const topos_batch_packet* packets, uint32_t count, uint16_t* const* planes_out,
                              const size_t* strides, topos_frame_output* infos)
{
    if (count == 0u) { return TC_OK; }
    if (packets == NULL || infos == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "batch args NULL"); }
    tc_view* views = (tc_view*)malloc(sizeof(tc_view) * count);
    tc_dec_req* reqs = (tc_dec_req*)calloc(count, sizeof(tc_dec_req));
    if (views == NULL || reqs == NULL) { free(views); free(reqs); return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch"); }
    int32_t rc = TC_OK;
    for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
        rc = dec_prepare(packets[i].data, packets[i].size, &infos[i], &views[i]);
    }
    if (rc != TC_OK || planes_out == NULL) { free(views); free(reqs); return rc; }
    uint32_t nrun = 0;
    int32_t first = TC_OK;
    tc_dec_req* run = (tc_dec_req*)calloc(count, sizeof(tc_dec_req));
    uint32_t* map = (uint32_t*)calloc(count, sizeof(uint32_t));
    for (uint32_t i = 0; i < count; ++i) {
        tc_dec_req* r = &reqs[i];
        r->data = packets[i].data; r->size = packets[i].size; r->view = &views[i]; r->info = &infos[i];
        r->planes = planes_out + (size_t)i * TC_FRAME_MAX_PLANES;
        r->strides = strides ? strides + (size_t)i * TC_FRAME_MAX_PLANES : NULL;
        r->rc = dec_supported(&views[i]);
        for (uint32_t p = 0; r->rc == TC_OK && p < views[i].fh.plane_count; ++p) {
            if (r->planes[p] == NULL) { r->rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "planes_out[%u][%u] == NULL", i, p); }
            else if (r->strides && r->strides[p] != 0u && r->strides[p] < views[i].fh.pvw[p]) {
                r->rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "frame %u plane %u stride too small", i, p);
            }
        }
        if (r->rc == TC_OK) { map[nrun] = i; run[nrun++] = *r; }
    }
    if (nrun > 0u) {
        tc_ectx* ec = NULL;
        if ((rc = tc_engine_acquire(&ec)) == TC_OK) {
            rc = tc_engine_decode(ec, run, nrun);
            tc_engine_release(ec);
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]].rc = rc < TC_OK ? rc : run[k].rc; }
        } else {
            for (uint32_t k = 0; k < nrun; ++k) { reqs[map[k]].rc = rc; }
        }
    }
    for (uint32_t i = 0; i < count; ++i) {
        if (first == TC_OK && reqs[i].rc != TC_OK) { first = reqs[i].rc; }
    }
    free(run); free(map); free(views); free(reqs);
    return first;
}

int32_t tc_frame_plane_geometry(const topos_frame_output* info, uint32_t plane, uint32_t* width, uint32_t* height)
{
    if (info == NULL || width == NULL || height == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane geometry args NULL");
    }
    if (plane >= info->plane_count) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane %u >= plane_count %u", plane, (unsigned)info->plane_count);
    }
    if (info->pixel_format == 3u) {
        *width = ((uint32_t)info->visible_width + 1u) / 2u;
        *height = ((uint32_t)info->visible_height + 1u) / 2u;
        return TC_OK;
    }
    const uint32_t cw = info->pixel_format == 0u ? ((uint32_t)info->visible_width + 1u) / 2u : info->visible_width;
    *width = (plane == 1u || plane == 2u) ? cw : info->visible_width;
    *height = info->visible_height;
    return TC_OK;
}

/* ------------------------------------------------------------ contexts */

struct tc_decoder { uint32_t workers; };

int32_t tc_decoder_create(const topos_decoder_config* cfg, tc_decoder** out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out == NULL"); }
    if (cfg != NULL && ((cfg->struct_size != 0u && cfg->struct_size != (uint32_t)sizeof(topos_decoder_config)) ||
                        (cfg->abi_version != 0u && cfg->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION))) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder config struct_size/abi_version mismatch");
    }
    tc_decoder* d = (tc_decoder*)calloc(1, sizeof(tc_decoder));
    if (d == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "decoder context"); }
    d->workers = cfg ? cfg->max_slice_workers : 0u;
    *out = d;
    return TC_OK;
}

void tc_decoder_destroy(tc_decoder* dec) { free(dec); }

int32_t tc_decoder_prepare(tc_decoder* dec, const uint8_t* packet, size_t size, topos_frame_output* info)
{
    if (dec == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder == NULL"); }
    return tc_frame_decode(packet, size, NULL, NULL, info);
}

static int32_t views_to_planes(const topos_plane_view* out, uint32_t n, uint16_t** planes, size_t* strides)
{
    for (uint32_t p = 0; p < n; ++p) {
        if ((out[p].struct_size != 0u && out[p].struct_size != (uint32_t)sizeof(topos_plane_view)) ||
            (out[p].abi_version != 0u && out[p].abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION)) {
            return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane view %u struct_size/abi_version mismatch", p);
        }
        planes[p] = out[p].pixels;
        strides[p] = out[p].stride;
    }
    return TC_OK;
}

int32_t tc_decoder_decode(tc_decoder* dec, const uint8_t* packet, size_t size,
                          const topos_plane_view out[TC_FRAME_MAX_PLANES], topos_frame_output* info)
{
    if (dec == NULL || out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder/out == NULL"); }
    uint16_t* planes[TC_FRAME_MAX_PLANES] = { 0 };
    size_t strides[TC_FRAME_MAX_PLANES] = { 0 };
    topos_frame_output tmp;
    int32_t rc = tc_frame_decode(packet, size, NULL, NULL, info ? info : &tmp);
    if (rc != TC_OK) { return rc; }
    const uint32_t np = (info ? info : &tmp)->plane_count;
    if ((rc = views_to_planes(out, np, planes, strides)) != TC_OK) { return rc; }
    return tc_frame_decode(packet, size, planes, strides, info ? info : &tmp);
}

int32_t tc_decoder_decode_batch(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                                const topos_plane_view* out, topos_frame_output* infos)
{
    if (dec == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder == NULL"); }
    if (out == NULL) { return tc_frame_decode_batch(packets, count, NULL, NULL, infos); }
    uint16_t** planes = (uint16_t**)calloc((size_t)count * TC_FRAME_MAX_PLANES + 1u, sizeof(uint16_t*));
    size_t* strides = (size_t*)calloc((size_t)count * TC_FRAME_MAX_PLANES + 1u, sizeof(size_t));
    if (planes == NULL || strides == NULL) { free(planes); free(strides); return tc_fail(TC_ERR_OUT_OF_MEMORY, "batch"); }
    int32_t rc = TC_OK;
    for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
        rc = views_to_planes(out + (size_t)i * TC_FRAME_MAX_PLANES, TC_FRAME_MAX_PLANES,
                             planes + (size_t)i * TC_FRAME_MAX_PLANES, strides + (size_t)i * TC_FRAME_MAX_PLANES);
    }
    if (rc == TC_OK) { rc = tc_frame_decode_batch(packets, count, planes, strides, infos); }
    free(planes);
    free(strides);
    return rc;
}
