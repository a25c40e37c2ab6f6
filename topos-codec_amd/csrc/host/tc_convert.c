/* Packed RGB(A) -> planar conversion and planar 4:4:4 -> 4:2:2 (reference
 * native/include/topos_codec.h:1016-1102, native/src/common/color_convert.c:
 * 353-449 and planar_convert.c:60-150): argument rules, derived constants and
 * error codes on the host, the pixels on the GPU. The host-buffer calls stage
 * the source into HBM, convert, and copy the planes back; tcx_*_dev take
 * device addresses in the same parameter structs. */
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

#define CVT_MAX_DIM 32768u

static uint32_t cvt_smp(uint32_t format) { return format >= (uint32_t)TOPOS_CVT_BGR48 ? 2u : 1u; }

static uint32_t cvt_nch(uint32_t format)
{
    return (format == TOPOS_CVT_BGRA32 || format == TOPOS_CVT_RGBA32 || format == TOPOS_CVT_BGRA64 ||
            format == TOPOS_CVT_RGBA64)
               ? 4u
               : 3u;
}

/* validation and constants in the reference's order (color_convert.c:353-427) */
static int32_t cvt_plan(const topos_cvt_params* p, tcg_cvt* k)
{
    if (p == NULL || p->struct_size != (uint32_t)sizeof(topos_cvt_params)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "cvt: params NULL or struct_size wrong");
    }
    if (p->format > (uint32_t)TOPOS_CVT_RGBA64) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "cvt: format %u", p->format); }
    if (p->out_mode > TOPOS_CVT_OUT_GBR || (p->bit_depth != 10u && p->bit_depth != 12u) || p->width == 0u ||
        p->height == 0u || p->width > CVT_MAX_DIM || p->height > CVT_MAX_DIM || p->src == NULL || p->y_out == NULL ||
        p->u_out == NULL || p->v_out == NULL || p->full_range > 1u || p->alpha_opaque > 1u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "cvt: argument out of range or NULL pointer");
    }
    memset(k, 0, sizeof(*k));
    k->mode = p->out_mode;
    k->smp = cvt_smp(p->format);
    k->nch = cvt_nch(p->format);
    k->width = p->width;
    k->height = p->height;
    k->cw = (p->width + 1u) / 2u;
    k->opaque = p->alpha_opaque;
    k->alpha_shift = p->alpha_shift > 15u ? 15u : p->alpha_shift;
    k->a_nmax = (float)(0xFFFFu >> k->alpha_shift);
    k->norm_div = k->smp == 2u ? 65535.0f : 255.0f;
    k->norm_rcp = 1.0f / k->norm_div;
    if (p->format == TOPOS_CVT_BGR24 || p->format == TOPOS_CVT_BGRA32 || p->format == TOPOS_CVT_BGR48 ||
        p->format == TOPOS_CVT_BGRA64) {
        k->ro = 2u; k->go = 1u; k->bo = 0u; k->ao = 3u;
    } else {
        k->ro = 0u; k->go = 1u; k->bo = 2u; k->ao = 3u;
    }
    k->y_st = p->y_stride ? p->y_stride : p->width;
    if (k->mode == TOPOS_CVT_OUT_YUV422) {
        k->u_st = p->u_stride ? p->u_stride : k->cw;
        k->v_st = p->v_stride ? p->v_stride : k->cw;
    } else {
        k->u_st = p->u_stride ? p->u_stride : p->width;
        k->v_st = p->v_stride ? p->v_stride : p->width;
    }
    k->a_st = p->a_stride ? p->a_stride : p->width;
    k->src_stride = p->src_stride;
    if (k->src_stride == 0u) {
        k->src_stride = p->width * k->nch * k->smp;
    } else if (k->src_stride < p->width * k->nch * k->smp) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "cvt: src_stride below the row width");
    }
    /* coefficients: double expressions rounded once to float */
    k->scale = (float)((1u << p->bit_depth) - 1u);
    k->kr = (float)p->kr;
    k->kb = (float)p->kb;
    if (k->mode != TOPOS_CVT_OUT_GBR) {
        const double krd = p->kr, kbd = p->kb;
        k->kf = (float)(1.0 - krd - kbd);
        k->du = (float)(2.0 * (1.0 - kbd));
        k->dv = (float)(2.0 * (1.0 - krd));
        k->rdu = 1.0 / (double)k->du;
        k->rdv = 1.0 / (double)k->dv;
        if (p->full_range) {
            k->y_mult = k->scale; k->y_off = 0.0f;
            k->c_mult = k->scale; k->c_off = 0.0f;
        } else {
            const uint32_t sh = p->bit_depth - 8u;
            k->y_mult = (float)(219u << sh); k->y_off = (float)(16u << sh);
            k->c_mult = (float)(224u << sh); k->c_off = (float)(16u << sh);
        }
    }
    return TC_OK;
}

static size_t al256(size_t v) { return (v + 255u) & ~(size_t)255u; }

int32_t tc_convert_packed_rgb(const topos_cvt_params* params)
{
    tcg_cvt k;
    int32_t rc = cvt_plan(params, &k);
    if (rc != TC_OK) { return rc; }
    const size_t row = (size_t)k.width * k.nch * k.smp;
    const size_t in_pitch = al256(row);
    const uint32_t cwid = k.mode == TOPOS_CVT_OUT_YUV422 ? k.cw : k.width;
    /* staged output rows start on 256-byte boundaries (the kernel's wide stores) */
    const uint32_t yp = (uint32_t)(al256((size_t)k.width * 2u) / 2u), cp = (uint32_t)(al256((size_t)cwid * 2u) / 2u);
    const size_t yb = (size_t)yp * k.height * 2u, cb = (size_t)cp * k.height * 2u;
    const int has_a = params->a_out != NULL;
    tc_ectx* c = NULL;
    if ((rc = tc_engine_acquire(&c)) != TC_OK) { return rc; }
    uint64_t din = 0, dout = 0;
    rc = tc_engine_staging(c, in_pitch * k.height, yb + 2u * cb + (has_a ? yb : 0u), &din, &dout);
    tcg_ctx* g = tc_engine_stream(c);
    if (rc == TC_OK && tcg_h2d_2d(g, din, in_pitch, params->src, k.src_stride, row, k.height) != 0) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "cvt: source upload: %s", tcg_last_error());
    }
    if (rc == TC_OK) {
        tcg_cvt d = k;
        d.src = din;
        d.src_stride = (uint32_t)in_pitch;
        d.y = dout;
        d.u = dout + yb;
        d.v = dout + yb + cb;
        d.a = has_a ? dout + yb + 2u * cb : 0u;
        d.y_st = yp;
        d.u_st = cp;
        d.v_st = cp;
        d.a_st = yp;
        if (tcg_convert_rgb(g, &d) != 0) { rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "cvt: %s", tcg_last_error()); }
        struct { uint16_t* dst; uint32_t st; uint64_t src; uint32_t pitch; uint32_t w; } outs[4] = {
            { params->y_out, k.y_st, d.y, yp, k.width }, { params->u_out, k.u_st, d.u, cp, cwid },
            { params->v_out, k.v_st, d.v, cp, cwid }, { params->a_out, k.a_st, d.a, yp, k.width } };
        for (int i = 0; i < (has_a ? 4 : 3) && rc == TC_OK; ++i) {
            if (tcg_d2h_2d(g, outs[i].dst, (size_t)outs[i].st * 2u, outs[i].src, (size_t)outs[i].pitch * 2u,
                           (size_t)outs[i].w * 2u, k.height) != 0) {
                rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "cvt: plane download: %s", tcg_last_error());
            }
        }
    }
    tc_engine_release(c);
    return rc;
}

int32_t tcx_convert_packed_rgb_dev(const topos_cvt_params* params)
{
    tcg_cvt k;
    int32_t rc = cvt_plan(params, &k);
    if (rc != TC_OK) { return rc; }
    k.src = (uint64_t)(uintptr_t)params->src;
    k.y = (uint64_t)(uintptr_t)params->y_out;
    k.u = (uint64_t)(uintptr_t)params->u_out;
    k.v = (uint64_t)(uintptr_t)params->v_out;
    k.a = (uint64_t)(uintptr_t)params->a_out;
    tc_ectx* c = NULL;
    if ((rc = tc_engine_acquire(&c)) != TC_OK) { return rc; }
    tcg_ctx* g = tc_engine_stream(c);
    if (tcg_convert_rgb(g, &k) != 0 || tcg_sync(g) != 0) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "cvt: %s", tcg_last_error());
    } else {
        tc_engine_record_convert(c);
    }
    tc_engine_release(c);
    return rc;
}

/* planar_convert.c:60-104 */
static int32_t p422_plan(const topos_planar_444_to_422_params* p, tcg_p422* k)
{
    if (p == NULL || p->struct_size != (uint32_t)sizeof(topos_planar_444_to_422_params) || p->width == 0u ||
        p->height == 0u || p->width > CVT_MAX_DIM || p->height > CVT_MAX_DIM ||
        (p->bit_depth != 10u && p->bit_depth != 12u && p->bit_depth != 16u) || p->u_in == NULL || p->v_in == NULL ||
        p->u_out == NULL || p->v_out == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "planar 444->422: argument out of range or NULL pointer");
    }
    memset(k, 0, sizeof(*k));
    k->width = p->width;
    k->height = p->height;
    k->cw = (p->width + 1u) / 2u;
    k->uis = p->u_in_stride ? p->u_in_stride : p->width;
    k->vis = p->v_in_stride ? p->v_in_stride : p->width;
    k->uos = p->u_out_stride ? p->u_out_stride : k->cw;
    k->vos = p->v_out_stride ? p->v_out_stride : k->cw;
    if (k->uis < p->width || k->vis < p->width || k->uos < k->cw || k->vos < k->cw) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "planar 444->422: stride below the plane width");
    }
    return TC_OK;
}

int32_t tc_convert_yuv444_to_422(const topos_planar_444_to_422_params* params)
{
    tcg_p422 k;
    int32_t rc = p422_plan(params, &k);
    if (rc != TC_OK) { return rc; }
    /* staged rows start on 256-byte boundaries (the kernel's vector accesses) */
    const size_t ip = al256((size_t)k.width * 2u), op = al256((size_t)k.cw * 2u);
    const size_t ib = ip * k.height, ob = op * k.height;
    tc_ectx* c = NULL;
    if ((rc = tc_engine_acquire(&c)) != TC_OK) { return rc; }
    uint64_t din = 0, dout = 0;
    rc = tc_engine_staging(c, 2u * ib, 2u * ob, &din, &dout);
    tcg_ctx* g = tc_engine_stream(c);
    tcg_p422 d = k;
    d.u_in = din; d.v_in = din + ib; d.u_out = dout; d.v_out = dout + ob;
    d.uis = ip / 2u; d.vis = ip / 2u; d.uos = op / 2u; d.vos = op / 2u;
    if (rc == TC_OK && (tcg_h2d_2d(g, d.u_in, ip, params->u_in, (size_t)k.uis * 2u,
                                   (size_t)k.width * 2u, k.height) != 0 ||
                        tcg_h2d_2d(g, d.v_in, ip, params->v_in, (size_t)k.vis * 2u,
                                   (size_t)k.width * 2u, k.height) != 0)) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "planar 444->422: upload: %s", tcg_last_error());
    }
    if (rc == TC_OK && tcg_convert_444_422(g, &d) != 0) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "planar 444->422: %s", tcg_last_error());
    }
    if (rc == TC_OK && (tcg_d2h_2d(g, params->u_out, (size_t)k.uos * 2u, d.u_out, op,
                                   (size_t)k.cw * 2u, k.height) != 0 ||
                        tcg_d2h_2d(g, params->v_out, (size_t)k.vos * 2u, d.v_out, op,
                                   (size_t)k.cw * 2u, k.height) != 0)) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "planar 444->422: download: %s", tcg_last_error());
    }
    tc_engine_release(c);
    return rc;
}

int32_t tcx_convert_yuv444_to_422_dev(const topos_planar_444_to_422_params* params)
{
    tcg_p422 k;
    int32_t rc = p422_plan(params, &k);
    if (rc != TC_OK) { return rc; }
    k.u_in = (uint64_t)(uintptr_t)params->u_in;
    k.v_in = (uint64_t)(uintptr_t)params->v_in;
    k.u_out = (uint64_t)(uintptr_t)params->u_out;
    k.v_out = (uint64_t)(uintptr_t)params->v_out;
    tc_ectx* c = NULL;
    if ((rc = tc_engine_acquire(&c)) != TC_OK) { return rc; }
    tcg_ctx* g = tc_engine_stream(c);
    if (tcg_convert_444_422(g, &k) != 0 || tcg_sync(g) != 0) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "planar 444->422: %s", tcg_last_error());
    } else {
        tc_engine_record_convert(c);
    }
    tc_engine_release(c);
    return rc;
}

float tcx_last_convert_ms(void) { return tc_engine_last_convert_ms(); }
