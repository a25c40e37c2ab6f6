/* Frozen synthetic frame generators for the benchmark workloads.
 *
 * Every sample is a pure function of (seed, plane tag, geometry): a
 * splitmix64 stream drives uniforms in [0,1) and Box-Muller normals in f64,
 * sinusoids are evaluated separably (sin(a+b) = sin a cos b + cos a sin b
 * over 1-D tables) and the field is rounded with rint() then clipped.
 *
 *  - video luma  : 512 + 6 sinusoids (1..16 cycles/frame, amplitude 20..120)
 *                  + 12 axis-aligned rectangles of +/-150 + N(0, 8^2),
 *                  clipped to [64, 940];
 *  - video chroma: 512 + 2 sinusoids (amplitude 60) + N(0, 4^2), [64, 960];
 *  - 12-bit GBR  : luma field per plane rescaled 256 + 4*(v-64) -> [256, 3760];
 *  - alpha       : 20 ellipses with a 4-pixel linear ramp, [0, 65535];
 *  - RAW Bayer   : luma field mapped to linear 12-bit, Poisson-Gaussian noise
 *                  (gain 2, read noise sigma 3), [0, 4095], RGGB split into
 *                  R / Gr / Gb / B phase planes.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s; int have; double spare; } tsyn_rng;

static uint64_t rng_next(tsyn_rng* r)
{
    uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double rng_u01(tsyn_rng* r) { return (double)(rng_next(r) >> 11) * 0x1.0p-53; }

static double rng_gauss(tsyn_rng* r)
{
    if (r->have) { r->have = 0; return r->spare; }
    const double u1 = 1.0 - rng_u01(r);
    const double u2 = rng_u01(r);
    const double rad = sqrt(-2.0 * log(u1));
    const double th = 6.283185307179586 * u2;
    r->spare = rad * sin(th);
    r->have = 1;
    return rad * cos(th);
}

static void rng_seed(tsyn_rng* r, uint64_t seed, uint32_t tag)
{
    r->s = seed * 0x100000001B3ull + (uint64_t)tag * 0x9E3779B97F4A7C15ull + 0x5DEECE66Dull;
    r->have = 0;
    r->spare = 0.0;
    (void)rng_next(r);
}

typedef struct {
    double base;
    int nsin;
    double amp_lo, amp_hi;
    int nrect;
    double rect_off;
    double sigma;
    double lo, hi;
} tsyn_field;

/* Writes a w*h double field (before rounding) into f. Returns 0 on success. */
static int field_eval(const tsyn_field* P, tsyn_rng* r, uint32_t w, uint32_t h, double* f)
{
    double* tx = (double*)malloc(sizeof(double) * 2u * w);
    double* ty = (double*)malloc(sizeof(double) * 2u * h);
    if (tx == NULL || ty == NULL) { free(tx); free(ty); return -1; }
    for (size_t i = 0; i < (size_t)w * h; ++i) { f[i] = P->base; }
    for (int k = 0; k < P->nsin; ++k) {
        const double fx = 1.0 + 15.0 * rng_u01(r);
        const double fy = 1.0 + 15.0 * rng_u01(r);
        const double amp = P->amp_lo + (P->amp_hi - P->amp_lo) * rng_u01(r);
        const double phi = 6.283185307179586 * rng_u01(r);
        for (uint32_t x = 0; x < w; ++x) {
            const double a = 6.283185307179586 * fx * (double)x / (double)w;
            tx[2u * x] = sin(a);
            tx[2u * x + 1u] = cos(a);
        }
        for (uint32_t y = 0; y < h; ++y) {
            const double b = 6.283185307179586 * fy * (double)y / (double)h + phi;
            ty[2u * y] = sin(b);
            ty[2u * y + 1u] = cos(b);
        }
        for (uint32_t y = 0; y < h; ++y) {
            const double sy = ty[2u * y], cy = ty[2u * y + 1u];
            double* row = f + (size_t)y * w;
            for (uint32_t x = 0; x < w; ++x) {
                const double s = tx[2u * x] * cy + tx[2u * x + 1u] * sy;
                row[x] += amp * s;
            }
        }
    }
    for (int k = 0; k < P->nrect; ++k) {
        const uint32_t x0 = (uint32_t)(rng_u01(r) * w);
        const uint32_t x1 = x0 + 1u + (uint32_t)(rng_u01(r) * (double)(w - x0));
        const uint32_t y0 = (uint32_t)(rng_u01(r) * h);
        const uint32_t y1 = y0 + 1u + (uint32_t)(rng_u01(r) * (double)(h - y0));
        const double off = rng_u01(r) < 0.5 ? -P->rect_off : P->rect_off;
        for (uint32_t y = y0; y < y1 && y < h; ++y) {
            for (uint32_t x = x0; x < x1 && x < w; ++x) { f[(size_t)y * w + x] += off; }
        }
    }
    free(tx);
    free(ty);
    return 0;
}

static int plane_from_field(const tsyn_field* P, uint64_t seed, uint32_t tag,
                            uint32_t w, uint32_t h, uint16_t* out, size_t stride,
                            int scale4_12bit)
{
    tsyn_rng r;
    rng_seed(&r, seed, tag);
    double* f = (double*)malloc(sizeof(double) * (size_t)w * h);
    if (f == NULL) { return -1; }
    if (field_eval(P, &r, w, h, f) != 0) { free(f); return -1; }
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            double v = f[(size_t)y * w + x] + P->sigma * rng_gauss(&r);
            v = rint(v);
            if (v < P->lo) { v = P->lo; }
            if (v > P->hi) { v = P->hi; }
            uint32_t u = (uint32_t)v;
            if (scale4_12bit) { u = 256u + (u - 64u) * 4u; }
            out[(size_t)y * stride + x] = (uint16_t)u;
        }
    }
    free(f);
    return 0;
}

static const tsyn_field kLuma = { 512.0, 6, 20.0, 120.0, 12, 150.0, 8.0, 64.0, 940.0 };
static const tsyn_field kChroma = { 512.0, 2, 60.0, 60.0, 0, 0.0, 4.0, 64.0, 960.0 };

/* 10-bit luma-like plane (w x h). stride in elements, 0 = tight. */
int tsyn_luma10(uint64_t seed, uint32_t tag, uint32_t w, uint32_t h, uint16_t* out, size_t stride)
{
    return plane_from_field(&kLuma, seed, tag, w, h, out, stride ? stride : w, 0);
}

/* 10-bit chroma-like plane. */
int tsyn_chroma10(uint64_t seed, uint32_t tag, uint32_t w, uint32_t h, uint16_t* out, size_t stride)
{
    return plane_from_field(&kChroma, seed, tag, w, h, out, stride ? stride : w, 0);
}

/* 12-bit colour plane: the luma field rescaled to [256, 3760]. */
int tsyn_color12(uint64_t seed, uint32_t tag, uint32_t w, uint32_t h, uint16_t* out, size_t stride)
{
    return plane_from_field(&kLuma, seed, tag, w, h, out, stride ? stride : w, 1);
}

/* 16-bit alpha: 20 ellipses, 4-pixel linear ramp at the rim. */
int tsyn_alpha16(uint64_t seed, uint32_t tag, uint32_t w, uint32_t h, uint16_t* out, size_t stride)
{
    if (stride == 0) { stride = w; }
    tsyn_rng r;
    rng_seed(&r, seed, tag);
    double cx[20], cy[20], rx[20], ry[20];
    for (int k = 0; k < 20; ++k) {
        cx[k] = rng_u01(&r) * w;
        cy[k] = rng_u01(&r) * h;
        rx[k] = (0.02 + 0.2 * rng_u01(&r)) * w + 2.0;
        ry[k] = (0.02 + 0.2 * rng_u01(&r)) * h + 2.0;
    }
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            double cov = 0.0;
            for (int k = 0; k < 20; ++k) {
                const double dx = ((double)x + 0.5 - cx[k]) / rx[k];
                const double dy = ((double)y + 0.5 - cy[k]) / ry[k];
                const double d = sqrt(dx * dx + dy * dy);
                const double rmin = rx[k] < ry[k] ? rx[k] : ry[k];
                double c = (1.0 - d) * rmin / 4.0;
                if (c > 1.0) { c = 1.0; }
                if (c > cov) { cov = c; }
            }
            out[(size_t)y * stride + x] = (uint16_t)rint(65535.0 * cov);
        }
    }
    return 0;
}

/* RAW 12-bit RGGB mosaic of a (2*pw) x (2*ph) field split into four
 * pw x ph phase planes R, Gr, Gb, B (planes[0..3], tight). */
int tsyn_raw12(uint64_t seed, uint32_t tag, uint32_t pw, uint32_t ph, uint16_t* const planes[4])
{
    const uint32_t w = 2u * pw, h = 2u * ph;
    tsyn_rng r;
    rng_seed(&r, seed, tag);
    double* f = (double*)malloc(sizeof(double) * (size_t)w * h);
    if (f == NULL) { return -1; }
    if (field_eval(&kLuma, &r, w, h, f) != 0) { free(f); return -1; }
    const double gain = 2.0, read_sigma = 3.0;
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t x = 0; x < w; ++x) {
            double lin = (f[(size_t)y * w + x] - 64.0) * (4095.0 / 876.0);
            if (lin < 0.0) { lin = 0.0; }
            const double lambda = lin / gain;
            /* Poisson photon count, normal approximation (lambda is large). */
            double n = lambda + sqrt(lambda) * rng_gauss(&r);
            if (n < 0.0) { n = 0.0; }
            double v = rint(gain * rint(n) + read_sigma * rng_gauss(&r));
            if (v < 0.0) { v = 0.0; }
            if (v > 4095.0) { v = 4095.0; }
            const uint32_t phase = ((y & 1u) << 1) | (x & 1u);
            planes[phase][(size_t)(y >> 1) * pw + (x >> 1)] = (uint16_t)v;
        }
    }
    free(f);
    return 0;
}

/* One 4:2:2 10-bit video frame: Y w x h, U/V ceil(w/2) x h (tight). */
int tsyn_frame422(uint64_t seed, uint32_t w, uint32_t h, uint16_t* y, uint16_t* u, uint16_t* v)
{
    const uint32_t cw = (w + 1u) / 2u;
    if (tsyn_luma10(seed, 0u, w, h, y, 0) != 0) { return -1; }
    if (tsyn_chroma10(seed, 1u, cw, h, u, 0) != 0) { return -1; }
    return tsyn_chroma10(seed, 2u, cw, h, v, 0);
}

/* One 4:4:4 12-bit GBR frame plus 16-bit alpha (all w x h, tight). */
int tsyn_frame4444(uint64_t seed, uint32_t w, uint32_t h, uint16_t* g, uint16_t* b,
                   uint16_t* r, uint16_t* a)
{
    if (tsyn_color12(seed, 10u, w, h, g, 0) != 0) { return -1; }
    if (tsyn_color12(seed, 11u, w, h, b, 0) != 0) { return -1; }
    if (tsyn_color12(seed, 12u, w, h, r, 0) != 0) { return -1; }
    return a != NULL ? tsyn_alpha16(seed, 13u, w, h, a, 0) : 0;
}
