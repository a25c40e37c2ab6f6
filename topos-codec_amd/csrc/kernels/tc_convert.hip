// gfx950 conversion kernels: packed RGB(A) -> YUV 4:2:2 / 4:4:4 / GBR planes
// and planar 4:4:4 -> 4:2:2 chroma.
//
// Reference semantics (paths relative to the reference tree):
//   packed RGB(A)   native/src/common/color_convert.c:50-124 (per-pixel float32
//                   math, quantiser, 4:2:2 box), :236-290 (row loop)
//   planar 444->422 native/src/common/planar_convert.c:26-45
//
// Float32 throughout in the reference's operation order; this file is built
// with contraction off (no implicit FMA) and HIP's IEEE-correct single-precision
// division, so every rounding matches the host build; the sample
// normalisation uses an exact reciprocal form (cvt_norm). rintf rounds half to
// even (v_rndne_f32).
#include <hip/hip_runtime.h>

#include "tcg.h"

namespace {

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

// sample / norm_div, correctly rounded like the IEEE division, for integer
// samples 0..norm_div: q0 = c r (r the correctly rounded 1 / d) is within an
// ulp of c / d, the residual c - q0 d is exact in one fma, and q0 + residual r
// in one more fma rounds to the nearest float (Markstein's correction); equal
// to c / d for every 8-bit and 16-bit sample (checked exhaustively)
__device__ inline float cvt_norm(const tcg_cvt& k, uint32_t v)
{
#pragma clang fp contract(off)
    const float c = (float)v;
    const float q0 = c * k.norm_rcp;
    const float res = __builtin_fmaf(-q0, k.norm_div, c);
    return __builtin_fmaf(res, k.norm_rcp, q0);
}

// the pixel pair's channels (x0 and, when two, x0 + 1), normalised, in
// storage order: 4-channel rows with a 16-byte aligned base and stride are
// read with one aligned vector load per pair, other rows sample by sample
__device__ inline void cvt_pair(const tcg_cvt& k, const uint8_t* row, uint32_t x0, bool two, float c[2][4])
{
    if (k.vec && k.smp == 1u) {
        uint32_t w0, w1 = 0u;
        if (two) {
            const uint2 w = *(const uint2*)(row + x0 * 4u);
            w0 = w.x;
            w1 = w.y;
        } else {
            w0 = *(const uint32_t*)(row + x0 * 4u);
        }
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            c[0][j] = cvt_norm(k, (w0 >> (8u * j)) & 255u);
            c[1][j] = cvt_norm(k, (w1 >> (8u * j)) & 255u);
        }
    } else if (k.vec) {
        uint32_t w[4] = { 0u, 0u, 0u, 0u };
        if (two) {
            const uint4 v = *(const uint4*)(row + x0 * 8u);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            const uint2 v = *(const uint2*)(row + x0 * 8u);
            w[0] = v.x; w[1] = v.y;
        }
#pragma unroll
        for (uint32_t p = 0; p < 2u; ++p) {
            c[p][0] = cvt_norm(k, w[2u * p] & 0xFFFFu);
            c[p][1] = cvt_norm(k, w[2u * p] >> 16);
            c[p][2] = cvt_norm(k, w[2u * p + 1u] & 0xFFFFu);
            c[p][3] = cvt_norm(k, w[2u * p + 1u] >> 16);
        }
    } else {
#pragma unroll
        for (uint32_t p = 0; p < 2u; ++p) {
#pragma unroll
            for (uint32_t j = 0; j < 4u; ++j) {
                const bool live = (p == 0u || two) && j < k.nch;
                const uint32_t off = ((x0 + p) * k.nch + j) * k.smp;
                uint32_t v = 0u;
                if (live) { v = k.smp == 1u ? row[off] : ((uint32_t)row[off] | ((uint32_t)row[off + 1u] << 8)); }
                c[p][j] = cvt_norm(k, v);
            }
        }
    }
}

// channel i (0..3) of a pixel without dynamic register indexing
__device__ inline float cvt_ch(const float* c, uint32_t i)
{
    return i == 0u ? c[0] : (i == 1u ? c[1] : (i == 2u ? c[2] : c[3]));
}

__device__ inline uint16_t cvt_alpha(const tcg_cvt& k, const float* c)
{
#pragma clang fp contract(off)
    if (k.opaque || k.nch == 3u) { return (uint16_t)(0xFFFFu << k.alpha_shift); }
    const float af = cvt_ch(c, k.ao);
    float v = rintf(af * k.a_nmax);
    if (v < 0.0f) { v = 0.0f; }
    if (v > k.a_nmax) { v = k.a_nmax; }
    return (uint16_t)((uint32_t)v << k.alpha_shift);
}

// y = kr r + kf g, then + kb b (left to right); u = (b - y) / du, v = (r - y) / dv
__device__ inline void cvt_yuv(const tcg_cvt& k, const float* c, uint16_t& yo, uint16_t& uo, uint16_t& vo)
{
#pragma clang fp contract(off)
    const float rf = cvt_ch(c, k.ro);
    const float gf = cvt_ch(c, k.go);
    const float bf = cvt_ch(c, k.bo);
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
    float c[2][4];
    cvt_pair(k, row, x0, two, c);
    if (k.mode == 2u) {
#pragma unroll
        for (uint32_t p = 0; p < 2u; ++p) {
            if (p == 1u && !two) { break; }
            yo[x0 + p] = cvt_q(cvt_ch(c[p], k.go) * k.scale, k.scale);
            uo[x0 + p] = cvt_q(cvt_ch(c[p], k.bo) * k.scale, k.scale);
            vo[x0 + p] = cvt_q(cvt_ch(c[p], k.ro) * k.scale, k.scale);
        }
    } else if (k.mode == 1u) {
        cvt_yuv(k, c[0], yo[x0], uo[x0], vo[x0]);
        if (two) { cvt_yuv(k, c[1], yo[x0 + 1u], uo[x0 + 1u], vo[x0 + 1u]); }
    } else {
        uint16_t y0v, u0v, v0v;
        cvt_yuv(k, c[0], y0v, u0v, v0v);
        yo[x0] = y0v;
        if (two) {
            uint16_t y1v, u1v, v1v;
            cvt_yuv(k, c[1], y1v, u1v, v1v);
            yo[x0 + 1u] = y1v;
            uo[cx] = cvt_box2(u0v, u1v);
            vo[cx] = cvt_box2(v0v, v1v);
        } else {
            uo[cx] = u0v;
            vo[cx] = v0v;
        }
    }
    if (ao != nullptr) {
        ao[x0] = cvt_alpha(k, c[0]);
        if (two) { ao[x0 + 1u] = cvt_alpha(k, c[1]); }
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
    tcg_cvt kv = *k;
    // aligned vector loads: 4-channel rows whose base and stride are 16-byte aligned
    kv.vec = kv.nch == 4u && (kv.src & 15u) == 0u && (kv.src_stride & 15u) == 0u;
    hipLaunchKernelGGL(k_convert_rgb, dim3((kv.cw + 255u) / 256u, kv.height), dim3(256), 0, c->stream, kv);
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
