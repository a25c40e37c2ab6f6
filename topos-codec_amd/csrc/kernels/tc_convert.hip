// gfx950 conversion kernels: packed RGB(A) -> YUV 4:2:2 / 4:4:4 / GBR planes
// and planar 4:4:4 -> 4:2:2 chroma.
//
// Reference semantics (paths relative to the reference tree):
//   packed RGB(A)   native/src/common/color_convert.c:50-124 (per-pixel float32
//                   math, quantiser, 4:2:2 box), :236-290 (row loop)
//   planar 444->422 native/src/common/planar_convert.c:26-45
//
// Float32 throughout in the reference's operation order; this file is built
// with contraction off (no implicit FMA), so every rounding matches the host
// build; the divisions are exact reciprocal forms that round like the IEEE
// quotient (cvt_norm for the sample normalisation, cvt_div for the chroma
// scale). rintf rounds half to even (v_rndne_f32).
#include <hip/hip_runtime.h>

#include "tcg.h"

namespace {

// clip(rint(v), 0, scale): v is always finite here, so the clamp is one
// median of three
__device__ inline float cvt_qf(float v, float scale) { return __builtin_amdgcn_fmed3f(rintf(v), 0.0f, scale); }

__device__ inline uint16_t cvt_q(float v, float scale) { return (uint16_t)cvt_qf(v, scale); }

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

// alpha of a pixel whose alpha channel (normalised) is af
__device__ inline uint16_t cvt_alpha_f(const tcg_cvt& k, float af)
{
#pragma clang fp contract(off)
    if (k.opaque || k.nch == 3u) { return (uint16_t)(0xFFFFu << k.alpha_shift); }
    const float v = __builtin_amdgcn_fmed3f(rintf(af * k.a_nmax), 0.0f, k.a_nmax);
    return (uint16_t)((uint32_t)v << k.alpha_shift);
}

__device__ inline uint16_t cvt_alpha(const tcg_cvt& k, const float* c) { return cvt_alpha_f(k, cvt_ch(c, k.ao)); }

// The binary32 quotient a / d, correctly rounded, as the binary64 product
// a (1 / d) rounded once to binary32 (rd = 1 / d correctly rounded in
// binary64). The product is within 2^-52 (relative) of a / d. A quotient of
// two binary32 numbers is never a binary32 midpoint (a midpoint needs 25
// significant bits, a / d = m would make a = m d carry them), and one that is
// not a midpoint lies more than 2^-49 (relative) from every midpoint (a - m d
// is a nonzero multiple of 2^-24 ulp(a)), so the product rounds to the same
// binary32 number as the exact quotient.
__device__ inline float cvt_div(float a, double rd) { return (float)((double)a * rd); }

// y = kr r + kf g, then + kb b (left to right); u = (b - y) / du, v = (r - y) / dv;
// the quantised samples as (integral) floats
__device__ inline void cvt_yuv_qf(const tcg_cvt& k, float rf, float gf, float bf, float& yo, float& uo, float& vo)
{
#pragma clang fp contract(off)
    const float p0 = k.kr * rf;
    const float p1 = k.kf * gf;
    float y = p0 + p1;
    const float p2 = k.kb * bf;
    y = y + p2;
    const float u = cvt_div(bf - y, k.rdu);
    const float v = cvt_div(rf - y, k.rdv);
    const float ys = y * k.y_mult;
    yo = cvt_qf(ys + k.y_off, k.scale);
    const float us = (u + 0.5f) * k.c_mult;
    uo = cvt_qf(us + k.c_off, k.scale);
    const float vs = (v + 0.5f) * k.c_mult;
    vo = cvt_qf(vs + k.c_off, k.scale);
}

__device__ inline void cvt_yuv_f(const tcg_cvt& k, float rf, float gf, float bf, uint16_t& yo, uint16_t& uo,
                                 uint16_t& vo)
{
    float yq, uq, vq;
    cvt_yuv_qf(k, rf, gf, bf, yq, uq, vq);
    yo = (uint16_t)yq;
    uo = (uint16_t)uq;
    vo = (uint16_t)vq;
}

__device__ inline void cvt_yuv(const tcg_cvt& k, const float* c, uint16_t& yo, uint16_t& uo, uint16_t& vo)
{
    cvt_yuv_f(k, cvt_ch(c, k.ro), cvt_ch(c, k.go), cvt_ch(c, k.bo), yo, uo, vo);
}

// converts and stores the pixel pair x0 (and x0 + 1 when two) of one row
__device__ inline void cvt_pair_store(const tcg_cvt& k, const float (&c)[2][4], uint32_t x0, bool two,
                                      uint16_t* yo, uint16_t* uo, uint16_t* vo, uint16_t* ao)
{
#pragma clang fp contract(off)
    const uint32_t cx = x0 >> 1;
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

// output rows of source row y
struct CvtRows {
    uint16_t *y, *u, *v, *a;
};

__device__ inline CvtRows cvt_rows(const tcg_cvt& k, uint32_t y)
{
    CvtRows r;
    r.y = (uint16_t*)k.y + (size_t)y * k.y_st;
    r.u = (uint16_t*)k.u + (size_t)y * k.u_st;
    r.v = (uint16_t*)k.v + (size_t)y * k.v_st;
    r.a = k.a ? (uint16_t*)k.a + (size_t)y * k.a_st : nullptr;
    return r;
}

// one thread per horizontal pixel pair (x = 2 cx, 2 cx + 1) of row blockIdx.y:
// any base address and stride
__global__ __launch_bounds__(256) void k_convert_rgb(const tcg_cvt k)
{
    const uint32_t cx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (cx >= k.cw || y >= k.height) { return; }
    const uint8_t* row = (const uint8_t*)k.src + (size_t)y * k.src_stride;
    const CvtRows o = cvt_rows(k, y);
    const uint32_t x0 = cx * 2u;
    const bool two = x0 + 1u < k.width;
    float c[2][4];
    cvt_pair(k, row, x0, two, c);
    cvt_pair_store(k, c, x0, two, o.y, o.u, o.v, o.a);
}

typedef uint32_t cvt_u32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t cvt_u32x4 __attribute__((ext_vector_type(4)));
// global-memory views (scalar base + vector offset addressing)
typedef __attribute__((address_space(1))) const uint32_t cvt_gsrc;
typedef __attribute__((address_space(1))) const cvt_u32x4 cvt_gsrc4;
typedef __attribute__((address_space(1))) uint32_t cvt_gdst;
typedef __attribute__((address_space(1))) cvt_u32x2 cvt_gdst2;

// Pixels per thread of the aligned conversion kernel.
#ifndef CVT_NPX
#define CVT_NPX 4u
#endif

// n samples (16-bit integers, or integral floats in 0..65535; n = 2, 4 or 8)
// as one store of n / 2 dwords of 16-bit samples, aligned to its size. The
// stores are non-temporal: the planes are not read again by this launch
// (3840x2160 BGRA32 to 4:2:2 + alpha: 0.0175 ms against 0.0187 ms with
// ordinary stores; non-temporal source loads were slower, 0.0205 ms).
template <uint32_t N, typename T>
__device__ inline void cvt_stn(uint16_t* p, const T (&v)[N])
{
    uint32_t w[N / 2u];
#pragma unroll
    for (uint32_t i = 0; i < N / 2u; ++i) { w[i] = (uint32_t)v[2u * i] | (uint32_t)v[2u * i + 1u] << 16; }
    if constexpr (N == 2u) {
        __builtin_nontemporal_store(w[0], (cvt_gdst*)p);
    } else if constexpr (N == 4u) {
        __builtin_nontemporal_store(cvt_u32x2{ w[0], w[1] }, (cvt_gdst2*)p);
    } else {
        __builtin_nontemporal_store(cvt_u32x4{ w[0], w[1], w[2], w[3] }, (__attribute__((address_space(1))) cvt_u32x4*)p);
    }
}

// One thread per NPX pixels (x = NPX t .. NPX t + NPX - 1) of row blockIdx.y,
// for a source whose base and stride are 16-byte aligned and output rows that
// take 2 NPX-byte stores (NPX bytes for 4:2:2 chroma): the pixels' NCH x SMP
// x NPX source bytes arrive in vector loads (16 B per load where the offset
// allows it, else 4-byte aligned dword groups) and every output plane takes
// one store per thread. The last group of a row, when the width is not a
// multiple of NPX, goes pair by pair. Channel order: red at RO, green at 1,
// blue at BO = 2 - RO, alpha at 3.
template <uint32_t NCH, uint32_t SMP, uint32_t RO>
__global__ __launch_bounds__(256) void k_convert_rgb4(const tcg_cvt k)
{
#pragma clang fp contract(off)
    constexpr uint32_t NPX = CVT_NPX;
    constexpr uint32_t BO = 2u - RO;
    constexpr uint32_t NW = NCH * SMP * NPX / 4u;  // source dwords per thread
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * NPX;
    const uint32_t y = blockIdx.y;
    if (x0 >= k.width || y >= k.height) { return; }
    const uint8_t* row = (const uint8_t*)k.src + (size_t)y * k.src_stride;
    const CvtRows o = cvt_rows(k, y);
    if (x0 + NPX > k.width) {
        for (uint32_t x = x0; x < k.width; x += 2u) {
            const bool two = x + 1u < k.width;
            float c[2][4];
            cvt_pair(k, row, x, two, c);
            cvt_pair_store(k, c, x, two, o.y, o.u, o.v, o.a);
        }
        return;
    }
    cvt_gsrc* src = (cvt_gsrc*)(row + (size_t)x0 * (NCH * SMP));
    uint32_t w[NW];
    if constexpr (NW % 4u == 0u && (NCH * SMP * NPX) % 16u == 0u) {
#pragma unroll
        for (uint32_t i = 0; i < NW; i += 4u) {
            const cvt_u32x4 v = *(cvt_gsrc4*)(src + i);
            w[i] = v.x; w[i + 1u] = v.y; w[i + 2u] = v.z; w[i + 3u] = v.w;
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < NW; ++i) { w[i] = src[i]; }
    }
    // sample j of pixel i: sample index s = NCH i + j in storage order
    float c[NPX][4];
#pragma unroll
    for (uint32_t i = 0; i < NPX; ++i) {
#pragma unroll
        for (uint32_t j = 0; j < 4u; ++j) {
            const uint32_t s = NCH * i + j;
            uint32_t v = 0u;
            if (j < NCH) {
                v = SMP == 1u ? (w[s >> 2] >> (8u * (s & 3u))) & 255u : (w[s >> 1] >> (16u * (s & 1u))) & 0xFFFFu;
            }
            c[i][j] = cvt_norm(k, v);
        }
    }
    float yv[NPX], uv[NPX], vv[NPX];
    if (k.mode == 2u) {
#pragma unroll
        for (uint32_t i = 0; i < NPX; ++i) {
            yv[i] = cvt_qf(c[i][1] * k.scale, k.scale);
            uv[i] = cvt_qf(c[i][BO] * k.scale, k.scale);
            vv[i] = cvt_qf(c[i][RO] * k.scale, k.scale);
        }
    } else {
#pragma unroll
        for (uint32_t i = 0; i < NPX; ++i) { cvt_yuv_qf(k, c[i][RO], c[i][1], c[i][BO], yv[i], uv[i], vv[i]); }
    }
    cvt_stn<NPX>(o.y + x0, yv);
    if (k.mode == 0u) {
        // the 4:2:2 box on the quantised values (cvt_box2)
        float ub[NPX / 2u], vb[NPX / 2u];
#pragma unroll
        for (uint32_t i = 0; i < NPX / 2u; ++i) {
            ub[i] = rintf((uv[2u * i] + uv[2u * i + 1u]) * 0.5f);
            vb[i] = rintf((vv[2u * i] + vv[2u * i + 1u]) * 0.5f);
        }
        cvt_stn<NPX / 2u>(o.u + (x0 >> 1), ub);
        cvt_stn<NPX / 2u>(o.v + (x0 >> 1), vb);
    } else {
        cvt_stn<NPX>(o.u + x0, uv);
        cvt_stn<NPX>(o.v + x0, vv);
    }
    if (o.a != nullptr) {
        uint16_t av[NPX];
#pragma unroll
        for (uint32_t i = 0; i < NPX; ++i) { av[i] = cvt_alpha_f(k, c[i][3]); }
        cvt_stn<NPX>(o.a + x0, av);
    }
}

// rounded mean of horizontally adjacent samples 2x and 2x + 1 (the last
// sample repeated at an odd width)
__device__ inline uint16_t p422_px(const uint16_t* in, uint32_t x, uint32_t width)
{
    const uint32_t sx = x * 2u;
    const uint32_t sx1 = sx + 1u < width ? sx + 1u : sx;
    return (uint16_t)(((uint32_t)in[sx] + (uint32_t)in[sx1] + 1u) >> 1);
}

// one thread per output sample of row blockIdx.y: any base address and stride
__global__ __launch_bounds__(256) void k_convert_444_422(const tcg_p422 k)
{
    const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t y = blockIdx.y;
    if (x >= k.cw || y >= k.height) { return; }
    const uint16_t* us = (const uint16_t*)k.u_in + (size_t)y * k.uis;
    const uint16_t* vs = (const uint16_t*)k.v_in + (size_t)y * k.vis;
    ((uint16_t*)k.u_out)[(size_t)y * k.uos + x] = p422_px(us, x, k.width);
    ((uint16_t*)k.v_out)[(size_t)y * k.vos + x] = p422_px(vs, x, k.width);
}

// four output samples of one plane row from one 16-byte load of eight input
// samples: (a + b + 1) >> 1 of each 16-bit pair, per dword
__device__ inline cvt_u32x2 p422_four(const cvt_u32x4 w)
{
    const uint32_t m[4] = { w.x, w.y, w.z, w.w };
    uint32_t r[4];
#pragma unroll
    for (uint32_t i = 0; i < 4u; ++i) { r[i] = ((m[i] & 0xFFFFu) + (m[i] >> 16) + 1u) >> 1; }
    return cvt_u32x2{ r[0] | r[1] << 16, r[2] | r[3] << 16 };
}

// One thread per four output samples (x = 4 t .. 4 t + 3) of row blockIdx.y
// in both chroma planes, for input rows whose base and stride are 16-byte
// aligned and output rows that take 8-byte stores: one 16-byte load and one
// 8-byte non-temporal store per plane. The last group of a row, when it is
// short of eight input samples, goes sample by sample.
__global__ __launch_bounds__(256) void k_convert_444_422_4(const tcg_p422 k)
{
    const uint32_t x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4u;
    const uint32_t y = blockIdx.y;
    if (x0 >= k.cw || y >= k.height) { return; }
    const uint16_t* us = (const uint16_t*)k.u_in + (size_t)y * k.uis;
    const uint16_t* vs = (const uint16_t*)k.v_in + (size_t)y * k.vis;
    uint16_t* uo = (uint16_t*)k.u_out + (size_t)y * k.uos;
    uint16_t* vo = (uint16_t*)k.v_out + (size_t)y * k.vos;
    if (2u * x0 + 8u > k.width) {
        for (uint32_t x = x0; x < k.cw; ++x) {
            uo[x] = p422_px(us, x, k.width);
            vo[x] = p422_px(vs, x, k.width);
        }
        return;
    }
    const cvt_u32x4 wu = *(cvt_gsrc4*)(us + 2u * x0);
    const cvt_u32x4 wv = *(cvt_gsrc4*)(vs + 2u * x0);
    __builtin_nontemporal_store(p422_four(wu), (cvt_gdst2*)(uo + x0));
    __builtin_nontemporal_store(p422_four(wv), (cvt_gdst2*)(vo + x0));
}

}  // namespace

// the launch helpers need the context's stream (tc_dev.h's tcg_ctx layout)
#include "tc_dev.h"

extern "C" int32_t tcg_convert_rgb(tcg_ctx* c, const tcg_cvt* k)
{
    c->kind = 3;
    tcg_mark(c, 0);
    tcg_cvt kv = *k;
    // aligned vector loads: rows whose base and stride are 16-byte aligned
    const bool in16 = (kv.src & 15u) == 0u && (kv.src_stride & 15u) == 0u;
    kv.vec = kv.nch == 4u && in16;
    // CVT_NPX pixels per thread: 2 CVT_NPX-byte stores of Y, alpha and
    // full-width chroma, CVT_NPX-byte stores of 4:2:2 chroma
    const auto al = [](uint64_t base, uint32_t st, uint32_t bytes) {
        return (base & (bytes - 1u)) == 0u && ((uint64_t)st * 2u & (bytes - 1u)) == 0u;
    };
    const uint32_t cb = kv.mode == 0u ? CVT_NPX : 2u * CVT_NPX;
    const bool out_ok = al(kv.y, kv.y_st, 2u * CVT_NPX) && al(kv.u, kv.u_st, cb) && al(kv.v, kv.v_st, cb) &&
                        (kv.a == 0u || al(kv.a, kv.a_st, 2u * CVT_NPX));
    if (in16 && out_ok) {
        typedef void (*cvt_fn)(tcg_cvt);
        // [channels 4][16-bit samples][blue first]
        static const cvt_fn fns[2][2][2] = {
            { { k_convert_rgb4<3u, 1u, 0u>, k_convert_rgb4<3u, 1u, 2u> },
              { k_convert_rgb4<3u, 2u, 0u>, k_convert_rgb4<3u, 2u, 2u> } },
            { { k_convert_rgb4<4u, 1u, 0u>, k_convert_rgb4<4u, 1u, 2u> },
              { k_convert_rgb4<4u, 2u, 0u>, k_convert_rgb4<4u, 2u, 2u> } } };
        const cvt_fn fn = fns[kv.nch == 4u][kv.smp == 2u][kv.ro == 2u];
        hipLaunchKernelGGL(fn, dim3((kv.width + 256u * CVT_NPX - 1u) / (256u * CVT_NPX), kv.height), dim3(256), 0,
                           c->stream, kv);
    } else {
        hipLaunchKernelGGL(k_convert_rgb, dim3((kv.cw + 255u) / 256u, kv.height), dim3(256), 0, c->stream, kv);
    }
    tcg_mark(c, 1);
    return tcg_check("convert launch");
}

extern "C" int32_t tcg_convert_444_422(tcg_ctx* c, const tcg_p422* k)
{
    c->kind = 3;
    tcg_mark(c, 0);
    const bool in16 = ((k->u_in | k->v_in | k->uis * 2u | k->vis * 2u) & 15u) == 0u;
    const bool out8 = ((k->u_out | k->v_out | k->uos * 2u | k->vos * 2u) & 7u) == 0u;
    if (in16 && out8) {
        hipLaunchKernelGGL(k_convert_444_422_4, dim3((k->cw + 1023u) / 1024u, k->height), dim3(256), 0, c->stream, *k);
    } else {
        hipLaunchKernelGGL(k_convert_444_422, dim3((k->cw + 255u) / 256u, k->height), dim3(256), 0, c->stream, *k);
    }
    tcg_mark(c, 1);
    return tcg_check("444 to 422 launch");
}
