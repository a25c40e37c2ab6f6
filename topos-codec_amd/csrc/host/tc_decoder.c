/* Decoder contexts of include/topos_codec.h (reference
 * native/src/codec/decoder_ctx.c, declarations topos_codec.h:502-571).
 *
 * A context is a validated handle with a worker budget. The reference keeps
 * grow-only host scratch per context so its steady state allocates nothing;
 * here the decode runs on the GPU engine, whose device buffers, pinned staging
 * and launch graphs are grow-only per engine context, so repeated decodes of
 * one geometry reuse them the same way. Argument rules and their order are the
 * reference's: handle, info, packet, output views (exact struct_size /
 * abi_version and non-NULL pixels for every plane the frame outputs); a batch
 * rejects as a whole only on a structure failure, a bad request or an inter
 * frame, and rejects a frame alone for a bad view.
 */
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"

#define TC_DECODER_TAG 0x5444454Bu /* validity tag of a live context */

struct tc_decoder {
    uint32_t tag;
    _Atomic uint32_t max_slice_workers;
};

static int valid(const tc_decoder* dec) { return dec != NULL && dec->tag == TC_DECODER_TAG; }

static int32_t bad_handle(void) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder handle invalid"); }

int32_t tc_decoder_create(const topos_decoder_config* cfg, tc_decoder** out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out == NULL"); }
    *out = NULL;
    if (cfg != NULL && (cfg->struct_size != (uint32_t)sizeof(topos_decoder_config) ||
                        cfg->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decoder config struct_size/abi_version");
    }
    tc_decoder* d = (tc_decoder*)calloc(1, sizeof(tc_decoder));
    if (d == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "decoder context"); }
    atomic_init(&d->max_slice_workers, cfg != NULL ? cfg->max_slice_workers : 0u);
    d->tag = TC_DECODER_TAG;
    *out = d;
    return TC_OK;
}

void tc_decoder_destroy(tc_decoder* dec)
{
    if (!valid(dec)) { return; }
    dec->tag = 0u;
    free(dec);
}

int32_t tc_decoder_set_max_slice_workers(tc_decoder* dec, uint32_t workers)
{
    if (!valid(dec)) { return bad_handle(); }
    atomic_store_explicit(&dec->max_slice_workers, workers, memory_order_release);
    return TC_OK;
}

int32_t tc_decoder_prepare(tc_decoder* dec, const uint8_t* packet, size_t size, topos_frame_output* info)
{
    if (!valid(dec)) { return bad_handle(); }
    if (info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "info == NULL"); }
    if (packet == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet == NULL"); }
    /* structure parse + the header fields only: the rest of *info (its
     * struct_size / abi_version stamp included) is the caller's, as in
     * decoder_ctx.c:170-197 */
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    const int32_t rc = tc_parse_structure(packet, size, v);
    if (rc == TC_OK) { tc_fill_output_info(&v->fh, info); }
    free(v);
    return rc;
}

int32_t tc_decoder_prepare_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                   const topos_decode_request* request, topos_frame_output* info)
{
    if (!valid(dec)) { return bad_handle(); }
    if (info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "info == NULL"); }
    return tc_frame_decode_request(packet, size, request, NULL, NULL, info);
}

/* a plane view the frame writes: exact struct_size / abi_version, pixels */
static int32_t view_check(const topos_plane_view* v, uint32_t p)
{
    if (v->struct_size != (uint32_t)sizeof(topos_plane_view) || v->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane view %u struct_size/abi_version", p);
    }
    if (v->pixels == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "plane view %u pixels == NULL", p); }
    return TC_OK;
}

/* single-frame decode; with_request = 0 decodes at full resolution, else
 * through `request` (whose rules reject NULL) */
static int32_t decode_one(tc_decoder* dec, const uint8_t* packet, size_t size, int with_request,
                          const topos_decode_request* request, const topos_plane_view out[TC_FRAME_MAX_PLANES],
                          topos_frame_output* info)
{
    if (!valid(dec)) { return bad_handle(); }
    if (info == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "info == NULL"); }
    if (packet == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet == NULL"); }
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "out == NULL"); }
    topos_decode_request full;
    if (!with_request) {
        memset(&full, 0, sizeof(full));
        full.struct_size = (uint32_t)sizeof(full);
        full.abi_version = (uint32_t)TOPOS_CODEC_ABI_VERSION;
        full.mode = TC_DECODE_MODE_FULL;
        request = &full;
    }
    /* structure, stamped info and target geometry first (query mode) */
    int32_t rc = tc_frame_decode_request(packet, size, request, NULL, NULL, info);
    if (rc != TC_OK) { return rc; }
    uint16_t* planes[TC_FRAME_MAX_PLANES] = { NULL, NULL, NULL, NULL };
    size_t strides[TC_FRAME_MAX_PLANES] = { 0u, 0u, 0u, 0u };
    for (uint32_t p = 0; p < info->plane_count; ++p) {
        if ((rc = view_check(&out[p], p)) != TC_OK) { return rc; }
        planes[p] = out[p].pixels;
        strides[p] = out[p].stride;
    }
    return tc_frame_decode_request(packet, size, request, planes, strides, info);
}

int32_t tc_decoder_decode(tc_decoder* dec, const uint8_t* packet, size_t size,
                          const topos_plane_view out[TC_FRAME_MAX_PLANES], topos_frame_output* info)
{
    return decode_one(dec, packet, size, 0, NULL, out, info);
}

int32_t tc_decoder_decode_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                  const topos_decode_request* request, const topos_plane_view out[TC_FRAME_MAX_PLANES],
                                  topos_frame_output* info)
{
    return decode_one(dec, packet, size, 1, request, out, info);
}

/* count x 4 views -> flat plane pointers; a view that fails its checks leaves
 * a NULL plane, which the batch core rejects for that frame alone when the
 * frame outputs that plane (views past a frame's plane count are not read) */
static int32_t batch_views(const topos_plane_view* out, uint32_t count, uint16_t*** planes, size_t** strides)
{
    const size_t n = (size_t)count * TC_FRAME_MAX_PLANES;
    *planes = (uint16_t**)calloc(n, sizeof(uint16_t*));
    *strides = (size_t*)calloc(n, sizeof(size_t));
    if (*planes == NULL || *strides == NULL) {
        free(*planes);
        free(*strides);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "decoder batch views");
    }
    for (size_t k = 0; k < n; ++k) {
        const topos_plane_view* v = &out[k];
        const int ok = v->struct_size == (uint32_t)sizeof(topos_plane_view) &&
                       v->abi_version == (uint32_t)TOPOS_CODEC_ABI_VERSION;
        (*planes)[k] = ok ? v->pixels : NULL;
        (*strides)[k] = ok ? v->stride : 0u;
    }
    return TC_OK;
}

static int32_t decode_batch(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                            const topos_decode_request* request, const topos_plane_view* out,
                            topos_frame_output* infos)
{
    if (!valid(dec)) { return bad_handle(); }
    if (out == NULL) { return tc_decode_batch_run(packets, count, 0u, request, NULL, NULL, infos); }
    if (count == 0u) { return TC_OK; }
    uint16_t** planes = NULL;
    size_t* strides = NULL;
    int32_t rc = batch_views(out, count, &planes, &strides);
    if (rc != TC_OK) { return rc; }
    rc = tc_decode_batch_run(packets, count, 0u, request, planes, strides, infos);
    free(planes);
    free(strides);
    return rc;
}

int32_t tc_decoder_decode_batch(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                                const topos_plane_view* out, topos_frame_output* infos)
{
    return decode_batch(dec, packets, count, NULL, out, infos);
}

int32_t tc_decoder_decode_batch_request(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                                        const topos_decode_request* request, const topos_plane_view* out,
                                        topos_frame_output* infos)
{
    return decode_batch(dec, packets, count, request, out, infos);
}

/* surface checks of decoder_ctx.c:469-508 */
static int32_t surface_check(const topos_decode_surface* s)
{
    if (s == NULL || (s->memory_type != TC_DECODE_MEMORY_CPU && s->memory_type != TC_DECODE_MEMORY_HOST_VISIBLE_GPU)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode surface invalid");
    }
    if (s->struct_size != (uint32_t)sizeof(topos_decode_surface) ||
        s->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode surface struct_size/abi_version");
    }
    return TC_OK;
}

int32_t tc_decoder_decode_surface(tc_decoder* dec, const uint8_t* packet, size_t size,
                                  const topos_decode_surface* surface, topos_frame_output* info)
{
    const int32_t rc = surface_check(surface);
    if (rc != TC_OK) { return rc; }
    return tc_decoder_decode(dec, packet, size, surface->planes, info);
}

int32_t tc_decoder_decode_surface_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                          const topos_decode_request* request,
                                          const topos_decode_surface* surface, topos_frame_output* info)
{
    int32_t rc = surface_check(surface);
    if (rc != TC_OK) { return rc; }
    if (request == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "decode surface request invalid"); }
    if (request->memory_type != surface->memory_type) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "request/surface memory_type mismatch");
    }
    return tc_decoder_decode_request(dec, packet, size, request, surface->planes, info);
}
