/* ORACLE — test infrastructure only. CPU restatement of the packed RGB(A)
 * -> planar conversion and the planar 4:4:4 -> 4:2:2 chroma step, scalar,
 * one pixel at a time, used to check the GPU kernels.
 *
 * Follows native/src/common/color_convert.c: the per-pixel float32 math
 * (:88-124, cvt_yuv_pixel / cvt_alpha / cvt_q / cvt_box2), the row loop
 * (:236-290) and the constants (:353-427); native/src/common/
 * planar_convert.c:26-45 for 4:4:4 -> 4:2:2. Built with -ffp-contract=off so
 * no product becomes a fused multiply-add. */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

/* same byte layout as the public topos_cvt_params */
typedef struct orc_cvt_params {
    uint32_t struct_size, format, width, height, out_mode, bit_depth;
    double kr, kb;
    uint32_t full_range, alpha_shift, alpha_opaque;
    const void* src;
    uint32_t src_stride;
    uint16_t* y_out;
    uint32_t y_stride;
    uint16_t* u_out;
    uint32_t u_stride;
    uint16_t* v_out;
    uint32_t v_stride;
    uint16_t* a_out;
    uint32_t a_stride;
    uint32_t max_workers;
} orc_cvt_params;

typedef struct orc_p422_params {
    uint32_t struct_size, width, height, bit_depth;
    const uint16_t* u_in;
    size_t u_in_stride;
    const uint16_t* v_in;
    size_t v_in_stride;
    uint16_t* u_out;
    size_t u_out_stride;
    uint16_t* v_out;
    size_t v_out_stride;
    uint32_t max_workers;
} orc_p422_params;

typedef struct cvk {
    uint32_t smp, nch, ro, go, bo, ao, mode, shift, opaque;
    float scale, kr, kb, kf, du, dv, ym, yo, cm, co, anmax, ndiv;
} cvk;

static float ld(const uint8_t* row, uint32_t ch, uint32_t x, const cvk* k)
{
    const uint32_t off = (x * k->nch + ch) * k->smp;
    if (k->smp == 1u) { return (float)row[off]; }
    return (float)((uint32_t)row[off] | ((uint32_t)row[off + 1u] << 8));
}

static uint16_t q(float v, float scale)
{
    float r = rintf(v);
    if (r < 0.0f) { r = 0.0f; }
    if (r > scale) { r = scale; }
    return (uint16_t)r;
}

static uint16_t box2(uint16_t a, uint16_t b) { return (uint16_t)rintf(((float)a + (float)b) * 0.5f); }

static uint16_t alpha(const cvk* k, const uint8_t* row, uint32_t x)
{
    if (k->opaque || k->nch == 3u) { return (uint16_t)(0xFFFFu << k->shift); }
    const float af = ld(row, k->ao, x, k) / k->ndiv;
    float v = rintf(af * k->anmax);
    if (v < 0.0f) { v = 0.0f; }
    if (v > k->anmax) { v = k->anmax; }
    return (uint16_t)((uint32_t)v << k->shift);
}

static void yuv(const cvk* k, const uint8_t* row, uint32_t x, uint16_t* yo, uint16_t* uo, uint16_t* vo)
{
    const float rf = ld(row, k->ro, x, k) / k->ndiv;
    const float gf = ld(row, k->go, x, k) / k->ndiv;
    const float bf = ld(row, k->bo, x, k) / k->ndiv;
    float y = k->kr * rf + k->kf * gf;
    y = y + k->kb * bf;
    const float u = (bf - y) / k->du;
    const float v = (rf - y) / k->dv;
    *yo = q(y * k->ym + k->yo, k->scale);
    *uo = q((u + 0.5f) * k->cm + k->co, k->scale);
    *vo = q((v + 0.5f) * k->cm + k->co, k->scale);
}

/* returns 0, or -1 on arguments the reference rejects */
int32_t orc_convert_packed_rgb(const orc_cvt_params* p)
{
    if (p == NULL || p->struct_size != (uint32_t)sizeof(orc_cvt_params) || p->format > 7u || p->out_mode > 2u ||
        (p->bit_depth != 10u && p->bit_depth != 12u) || p->width == 0u || p->height == 0u ||
        p->width > 32768u || p->height > 32768u || p->src == NULL || p->y_out == NULL || p->u_out == NULL ||
        p->v_out == NULL || p->full_range > 1u || p->alpha_opaque > 1u) {
        return -1;
    }
    cvk k;
    memset(&k, 0, sizeof(k));
    k.mode = p->out_mode;
    k.smp = p->format >= 4u ? 2u : 1u;
    k.nch = (p->format == 2u || p->format == 3u || p->format == 6u || p->format == 7u) ? 4u : 3u;
    k.opaque = p->alpha_opaque;
    k.shift = p->alpha_shift > 15u ? 15u : p->alpha_shift;
    k.anmax = (float)(0xFFFFu >> k.shift);
    k.ndiv = k.smp == 2u ? 65535.0f : 255.0f;
    if (p->format == 0u || p->format == 2u || p->format == 4u || p->format == 6u) {
        k.ro = 2u; k.go = 1u; k.bo = 0u; k.ao = 3u;
    } else {
        k.ro = 0u; k.go = 1u; k.bo = 2u; k.ao = 3u;
    }
    const uint32_t w = p->width, cw = (w + 1u) / 2u;
    const uint32_t yst = p->y_stride ? p->y_stride : w;
    const uint32_t ust = p->u_stride ? p->u_stride : (k.mode == 0u ? cw : w);
    const uint32_t vst = p->v_stride ? p->v_stride : (k.mode == 0u ? cw : w);
    const uint32_t ast = p->a_stride ? p->a_stride : w;
    uint32_t sst = p->src_stride;
    if (sst == 0u) { sst = w * k.nch * k.smp; }
    else if (sst < w * k.nch * k.smp) { return -1; }
    k.scale = (float)((1u << p->bit_depth) - 1u);
    k.kr = (float)p->kr;
    k.kb = (float)p->kb;
    if (k.mode != 2u) {
        k.kf = (float)(1.0 - p->kr - p->kb);
        k.du = (float)(2.0 * (1.0 - p->kb));
        k.dv = (float)(2.0 * (1.0 - p->kr));
        if (p->full_range) {
            k.ym = k.scale; k.yo = 0.0f; k.cm = k.scale; k.co = 0.0f;
        } else {
            const uint32_t sh = p->bit_depth - 8u;
            k.ym = (float)(219u << sh); k.yo = (float)(16u << sh);
            k.cm = (float)(224u << sh); k.co = (float)(16u << sh);
        }
    }
    for (uint32_t y = 0; y < p->height; ++y) {
        const uint8_t* row = (const uint8_t*)p->src + (size_t)y * sst;
        uint16_t* yo = p->y_out + (size_t)y * yst;
        uint16_t* uo = p->u_out + (size_t)y * ust;
        uint16_t* vo = p->v_out + (size_t)y * vst;
        uint16_t* ao = p->a_out ? p->a_out + (size_t)y * ast : NULL;
        if (k.mode == 2u) {
            for (uint32_t x = 0; x < w; ++x) {
                const float gf = ld(row, k.go, x, &k) / k.ndiv;
                const float bf = ld(row, k.bo, x, &k) / k.ndiv;
                const float rf = ld(row, k.ro, x, &k) / k.ndiv;
                yo[x] = q(gf * k.scale, k.scale);
                uo[x] = q(bf * k.scale, k.scale);
                vo[x] = q(rf * k.scale, k.scale);
                if (ao) { ao[x] = alpha(&k, row, x); }
            }
        } else if (k.mode == 1u) {
            for (uint32_t x = 0; x < w; ++x) {
                yuv(&k, row, x, &yo[x], &uo[x], &vo[x]);
                if (ao) { ao[x] = alpha(&k, row, x); }
            }
        } else {
            for (uint32_t cx = 0; cx < cw; ++cx) {
                const uint32_t x0 = 2u * cx;
                uint16_t y0, u0, v0, y1, u1, v1;
                yuv(&k, row, x0, &y0, &u0, &v0);
                yo[x0] = y0;
                if (x0 + 1u < w) {
                    yuv(&k, row, x0 + 1u, &y1, &u1, &v1);
                    yo[x0 + 1u] = y1;
                    uo[cx] = box2(u0, u1);
                    vo[cx] = box2(v0, v1);
                } else {
                    uo[cx] = u0;
                    vo[cx] = v0;
                }
                if (ao) {
                    ao[x0] = alpha(&k, row, x0);
                    if (x0 + 1u < w) { ao[x0 + 1u] = alpha(&k, row, x0 + 1u); }
                }
            }
        }
    }
    return 0;
}

int32_t orc_convert_444_422(const orc_p422_params* p)
{
    if (p == NULL || p->struct_size != (uint32_t)sizeof(orc_p422_params) || p->width == 0u || p->height == 0u ||
        p->width > 32768u || p->height > 32768u ||
        (p->bit_depth != 10u && p->bit_depth != 12u && p->bit_depth != 16u) || p->u_in == NULL || p->v_in == NULL ||
        p->u_out == NULL || p->v_out == NULL) {
        return -1;
    }
    const uint32_t cw = (p->width + 1u) / 2u;
    const size_t uis = p->u_in_stride ? p->u_in_stride : p->width, vis = p->v_in_stride ? p->v_in_stride : p->width;
    const size_t uos = p->u_out_stride ? p->u_out_stride : cw, vos = p->v_out_stride ? p->v_out_stride : cw;
    if (uis < p->width || vis < p->width || uos < cw || vos < cw) { return -1; }
    for (uint32_t y = 0; y < p->height; ++y) {
        for (uint32_t x = 0; x < cw; ++x) {
            const uint32_t sx = 2u * x, sx1 = sx + 1u < p->width ? sx + 1u : sx;
            p->u_out[y * uos + x] =
                (uint16_t)(((uint32_t)p->u_in[y * uis + sx] + (uint32_t)p->u_in[y * uis + sx1] + 1u) >> 1);
            p->v_out[y * vos + x] =
                (uint16_t)(((uint32_t)p->v_in[y * vis + sx] + (uint32_t)p->v_in[y * vis + sx1] + 1u) >> 1);
        }
    }
    return 0;
}
