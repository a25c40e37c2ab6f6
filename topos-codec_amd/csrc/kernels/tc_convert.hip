// gfx950 conversion kernels: packed RGB(A) -> YUV 4:2:2 / 4:4:4 / GBR planes
// and planar 4:4:4 -> 4:2:2 chroma.
//
// Reference semantics (paths relative to the reference tree):
//   packed RGB(A)   native/src/common/color_convert.c:50-124 (per-pixel float32
//                   math, quantiser, 4:2:2 box), :236-290 (row loop)
//   planar 444->422 native/src/common/planar_convert.c:26-45
//
// Float32 throughout in the reference's operation order; this file is built
// with contraction off (no FMA) and HIP's IEEE-correct single-precision
// division, so every rounding matches the host build. rintf rounds half to
// even (v_rndne_f32).
#include <hip/hip_runtime.h>

#include "tcg.h"

namespace {

__device__ inline float cvt_load(const uint8_t* row, uint32_t off, uint32_t smp)
{
    if (smp == 1u) { return (float)row[off]; }
    return (float)((uint32_t)row[off] | ((uint32_t)row[off + 1u] << 8));
}

// clip(rint(v), 0, scale)
__device__ inline uint16_t cvt_q(float v, float scale)
{
    float r = rintf(v);
    if (r < 0.0f) { r = 0.0f; }
    if (r > scale) { r = scale; }
    return (uint16_t)r;
}

// horizontal box of two quantised samples (integers or .5, exact in float32)
__device__ inline uint16_t cvt_box2(uint16_t a, uint16_t b)
{
#pragma clang fp contract(off)
    return (uint16_t)rintf(((float)a + (float)b) * 0.5f);
}

__device__ inline float cvt_px(const tcg_cvt& k, const uint8_t* row, uint32_t ch, uint32_t x)
{
    return cvt_load(row, (x * k.nch + ch) * k.smp, k.smp);
}

__device__ inline uint16_t cvt_alpha(const tcg_cvt& k, const uint8_t* row, uint32_t x)
{
#pragma clang fp contract(off)
    if (k.opaque || k.nch == 3u) { return (uint16_t)(0xFFFFu << k.alpha_shift); }
    const float af = cvt_px(k, row, k.ao, x) / k.norm_div;
    float v = rintf(af * k.a_nmax);
    if (v < 0.0f) { v = 0.0f; }
    if (v > k.a_nmax) { v = k.a_nmax; }
    return (uint16_t)((uint32_t)v << k.alpha_shift);
}

// y = kr r + kf g, then + kb b (left to right); u = (b - y) / du, v = (r - y) / dv
__device__ inline void cvt_yuv(const tcg_cvt& k, const uint8_t* row, uint32_t x, uint16_t& yo, uint16_t& uo,
                               uint16_t& vo)
{
#pragma clang fp contract(off)
    const float rf = cvt_px(k, row, k.ro, x) / k.norm_div;
    const float gf = cvt_px(k, row, k.go, x) / k.norm_div;
    const float bf = cvt_px(k, row, k.bo, x) / k.norm_div;
    const float p0 = k.kr * rf;
    const float p1 = k.kf * gf;
    float y = p0 + p1;
    const float p2 = k.kb * bf;
    y = y + p2;
    const float u = (bf - y) / k.du;
    const float v = (rf - y) / k.dv;
    const float ys = y * k.y_mult;
    yo = cvt_q(ys + k.y_off, k.scale);
    const float us = (u + 0.5f) * k.c_mult;
    uo = cvt_q(us + k.c_off, k.scale);
    const float vs = (v + 0.5f) * k.c_mult;
    vo = cvt_q(vs + k.c_off, k.scale);
}

// one thread per horizontal pixel pair (x = 2 cx, 2 cx + 1) of row blockIdx.y
__global__ __launch_bounds__(256) void k_convert_rgb(const tcg_cvt k)
{
#pragma clang fp contract(off)
    const uint32_t cx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (cx >= k.cw || y >= k.height) { return; }
    const uint8_t* row = (const uint8_t*)k.src + (size_t)y * k.src_stride;
    uint16_t* yo = (uint16_t*)k.y + (size_t)y * k.y_st;
    uint16_t* uo = (uint16_t*)k.u + (size_t)y * k.u_st;
    uint16_t* vo = (uint16_t*)k.v + (size_t)y * k.v_st;
    uint16_t* ao = k.a ? (uint16_t*)k.a + (size_t)y * k.a_st : nullptr;
    const uint32_t x0 = cx * 2u;
    const bool two = x0 + 1u < k.width;
    if (k.mode == 2u) {
        for (uint32_t x = x0; x < x0 + (two ? 2u : 1u); ++x) {
            const float gf = cvt_px(k, row, k.go, x) / k.norm_div;
            const float bf = cvt_px(k, row, k.bo, x) / k.norm_div;
            const float rf = cvt_px(k, row, k.ro, x) / k.norm_div;
            yo[x] = cvt_q(gf * k.scale, k.scale);
            uo[x] = cvt_q(bf * k.scale, k.scale);
            vo[x] = cvt_q(rf * k.scale, k.scale);
        }
    } else if (k.mode == 1u) {
        for (uint32_t x = x0; x < x0 + (two ? 2u : 1u); ++x) { cvt_yuv(k, row, x, yo[x], uo[x], vo[x]); }
    } else {
        uint16_t y0v, u0v, v0v;
        cvt_yuv(k, row, x0, y0v, u0v, v0v);
        yo[x0] = y0v;
        if (two) {
            uint16_t y1v, u1v, v1v;
            cvt_yuv(k, row, x0 + 1u, y1v, u1v, v1v);
            yo[x0 + 1u] = y1v;
            uo[cx] = cvt_box2(u0v, u1v);
            vo[cx] = cvt_box2(v0v, v1v);
        } else {
            uo[cx] = u0v;
            vo[cx] = v0v;
        }
    }
    if (ao != nullptr) {
        ao[x0] = cvt_alpha(k, row, x0);
        if (two) { ao[x0 + 1u] = cvt_alpha(k, row, x0 + 1u); }
    }
}

__global__ __launch_bounds__(256) void k_convert_444_422(const tcg_p422 k)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= k.cw || y >= k.height) { return; }
    const uint32_t sx = x * 2u;
    const uint32_t sx1 = sx + 1u < k.width ? sx + 1u : sx;
    const uint16_t* us = (const uint16_t*)k.u_in + (size_t)y * k.uis;
    const uint16_t* vs = (const uint16_t*)k.v_in + (size_t)y * k.vis;
    ((uint16_t*)k.u_out)[(size_t)y * k.uos + x] = (uint16_t)(((uint32_t)us[sx] + (uint32_t)us[sx1] + 1u) >> 1);
    ((uint16_t*)k.v_out)[(size_t)y * k.vos + x] = (uint16_t)(((uint32_t)vs[sx] + (uint32_t)vs[sx1] + 1u) >> 1);
}

}  // namespace

// the launch helpers need the context's stream (tc_dev.h's tcg_ctx layout)
#include "tc_dev.h"

extern "C" int32_t tcg_convert_rgb(tcg_ctx* c, const tcg_cvt* k)
{
    c->kind = 3;
    tcg_mark(c, 0);
    hipLaunchKernelGGL(k_convert_rgb, dim3((k->cw + 255u) / 256u, k->height), dim3(256), 0, c->stream, *k);
    tcg_mark(c, 1);
    return tcg_check("convert launch");
}

extern "C" int32_t tcg_convert_444_422(tcg_ctx* c, const tcg_p422* k)
{
    c->kind = 3;
    tcg_mark(c, 0);
    hipLaunchKernelGGL(k_convert_444_422, dim3((k->cw + 255u) / 256u, k->height), dim3(256), 0, c->stream, *k);
    tcg_mark(c, 1);
    return tcg_check("444 to 422 launch");
}
