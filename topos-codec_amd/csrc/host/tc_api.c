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

int32_t tc_dec_prepare(const uint8_t* data, size_t size, topos_frame_output* info, tc_view* v)
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

int32_t tc_dec_supported(const tc_view* v)
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
    int32_t rc = tc_dec_prepare(data, size, info, v);
    return rc == TC_OK ? tc_dec_supported(v) : rc;
}

int32_t tc_frame_decode(const uint8_t* data, size_t size, uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                        const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    int32_t rc = tc_dec_prepare(data, size, out_info, v);
    if (rc == TC_OK && planes_out != NULL) {
        rc = tc_dec_supported(v);
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

/* reference codec.c:9792-9798 (the batch contract at topos_codec.h:374-400) */
int32_t tc_frame_decode_batch(const topos_batch_packet* packets, uint32_t count, uint16_t* const* planes_out,
                              const size_t* strides, topos_frame_output* infos)
{
    return tc_decode_batch_run(packets, count, (uint32_t)TC_DECODE_SCALE_FULL, NULL, planes_out, strides, infos);
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
