/* ORACLE — test infrastructure only. CPU restatement of the target-size
 * encode (tc_frame_encode_sized): a deterministic qp search whose every probe
 * is a whole-frame encode of this oracle, then the packet at the chosen qp.
 *
 * Follows native/src/codec/codec.c:
 *   :7795-7890  entry: validation, search mode, fallback, qp_used
 *   :7578-7787  strict search (log-slope jumps + bracket interpolation,
 *               cross-call seed from the last two working points)
 *   :7554-7576  strict_slope_from_pts
 *   :7199-7265  legacy linear search (+8 / +1 / reclaim below 3/4 target)
 *   :7375-7395, :7404-7551  "r6" fast search and its slope fit
 *   :7303-7351  probe bookkeeping: budget 24, monotonicity violations
 *   :6137-6143  the 512 MiB F-cache rule (exact probes available or not)
 * The reference's exact probes count the bytes a whole-frame encode writes,
 * so a probe here is the oracle's encode at that qp. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc.h"

#define QP_MAX 95u
#define BUDGET 24u
#define FCACHE_MAX (512ull * 1024ull * 1024ull)

typedef struct hint_state {
    uint32_t qp, valid, bytes, prev, prev_bytes;
} hint_state;

static hint_state g_hint;                 /* the reference keeps it in its process-wide encoder cache */
static int32_t g_mode = -1;               /* -1: read TOPOS_SIZED_SEARCH once */
static int32_t g_hint_off, g_model_off;

void orc_sized_reset(void) { memset(&g_hint, 0, sizeof(g_hint)); }
void orc_sized_set_mode(int32_t mode) { g_mode = mode < 0 || mode > 2 ? 0 : mode; }
void orc_sized_get_state(uint32_t s[5])
{
    s[0] = g_hint.qp; s[1] = g_hint.valid; s[2] = g_hint.bytes; s[3] = g_hint.prev; s[4] = g_hint.prev_bytes;
}

void orc_sized_put_state(const uint32_t s[5])
{
    g_hint.qp = s[0]; g_hint.valid = s[1]; g_hint.bytes = s[2]; g_hint.prev = s[3]; g_hint.prev_bytes = s[4];
}

void orc_sized_set_hint(int32_t disable, int32_t model_disable)
{
    g_hint_off = disable != 0;
    g_model_off = model_disable != 0;
}

typedef struct frame_src {
    const orc_cfg* cfg;
    const uint16_t* const* planes;
    const size_t* strides;
    uint8_t* buf;
    size_t cap;
    uint32_t known[QP_MAX + 1];
    uint32_t bytes[QP_MAX + 1];
    uint32_t count;             /* probes + final encodes, as the reference counts them */
} frame_src;

/* whole-frame encode at qp into the scratch buffer; bytes of the packet */
static int32_t encode_qp(frame_src* f, uint32_t qp, uint32_t* bytes)
{
    orc_cfg c = *f->cfg;
    c.qp_base = (uint8_t)qp;
    size_t n = 0;
    const int32_t rc = orc_frame_encode(&c, f->planes, f->strides, f->buf, f->cap, &n);
    if (rc < 0) { return rc; }
    f->known[qp] = 1;
    f->bytes[qp] = (uint32_t)n;
    *bytes = (uint32_t)n;
    return ORC_OK;
}

typedef struct probes {
    frame_src* f;
    uint32_t qp[BUDGET + 1], b[BUDGET + 1], n;
    uint32_t budget;
    int violated;
    uint32_t last_qp;
    int has_last;
    uint32_t size;              /* bytes of the last probe */
} probes;

static int32_t probe(probes* s, uint32_t qp)
{
    if (s->budget == 0u) { s->violated = 1; return ORC_ELIMIT; }
    uint32_t b = 0;
    const int32_t rc = encode_qp(s->f, qp, &b);
    s->f->count++;
    s->budget--;
    s->last_qp = qp;
    s->has_last = 1;
    if (rc < 0) { return rc; }
    s->size = b;
    for (uint32_t i = 0; i < s->n; ++i) {
        if ((s->qp[i] < qp && s->b[i] < b) || (s->qp[i] > qp && s->b[i] > b)) { s->violated = 1; }
    }
    s->qp[s->n] = qp;
    s->b[s->n] = b;
    s->n++;
    return ORC_OK;
}

static int seen(const probes* s, uint32_t qp, uint32_t* b)
{
    for (uint32_t i = 0; i < s->n; ++i) {
        if (s->qp[i] == qp) { *b = s->b[i]; return 1; }
    }
    return 0;
}

/* newest pair (q rises, bytes fall) with a log2 ratio above 0.05 */
static double pair_slope(const probes* s, double dflt, double lo_clamp, double hi_clamp, int need_b0)
{
    for (uint32_t i = s->n; i-- > 1u;) {
        for (uint32_t j = i; j-- > 0u;) {
            const uint32_t q1 = s->qp[i], b1 = s->b[i], q0 = s->qp[j], b0 = s->b[j];
            if (q1 > q0 && b0 > b1 && (need_b0 ? b0 > 0u : b1 > 0u)) {
                const double dr = log2((double)b0 / (double)b1);
                if (dr > 0.05) {
                    double v = (double)(q1 - q0) / dr;
                    if (v < lo_clamp) { v = lo_clamp; }
                    if (v > hi_clamp) { v = hi_clamp; }
                    return v;
                }
            }
        }
    }
    return dflt;
}

static uint32_t clampu(uint32_t v, uint32_t lo, uint32_t hi) { return v < lo ? lo : v > hi ? hi : v; }

/* the working-point pair of the hint: (q_hi, b_hi) above (q_lo, b_lo) */
static int hint_pair(const hint_state* h, double* slope_out)
{
    const uint32_t q_hi = h->qp > h->prev ? h->qp : h->prev;
    const uint32_t q_lo = q_hi == h->qp ? h->prev : h->qp;
    const uint32_t b_hi = q_hi == h->qp ? h->bytes : h->prev_bytes;
    const uint32_t b_lo = q_hi == h->qp ? h->prev_bytes : h->bytes;
    if (q_hi > q_lo && b_lo > b_hi && b_hi > 0u) {
        *slope_out = (double)(q_hi - q_lo) / log2((double)b_lo / (double)b_hi);
        return 1;
    }
    return 0;
}

/* final packet at qp into the caller's buffer (capacity checked) */
static int32_t emit_final(frame_src* f, uint32_t qp, uint8_t* out, size_t cap, size_t* size_out)
{
    uint32_t b = 0;
    const int32_t rc = encode_qp(f, qp, &b);
    if (rc < 0) { return rc; }
    *size_out = b;
    if (b > cap) { return ORC_EBUF; }
    if (out != NULL) { memcpy(out, f->buf, b); }
    return ORC_OK;
}

static int32_t search_strict(frame_src* f, int exact, uint32_t target, uint32_t qp_min, uint32_t qp_max,
                             uint8_t* out, size_t cap, size_t* size_out, uint32_t* q_used, int* fallback)
{
    if (!exact) { *fallback = 1; return ORC_OK; }
    probes s;
    memset(&s, 0, sizeof(s));
    s.f = f;
    s.budget = BUDGET;
    uint32_t q0 = f->cfg->qp_base;
    if (!g_hint_off && g_hint.valid) {
        q0 = g_hint.qp;
        if (!g_model_off && g_hint.bytes > 0u && target > 0u) {
            const double dr = log2((double)g_hint.bytes / (double)target);
            double slope = 0.0;
            if (g_hint.prev_bytes > 0u && g_hint.prev != g_hint.qp) { (void)hint_pair(&g_hint, &slope); }
            if (slope > 0.0) {
                slope = slope < 0.25 ? 0.25 : slope > 32.0 ? 32.0 : slope;
            } else {
                slope = 8.0;
            }
            double dq = slope * dr;
            dq = dq > 32.0 ? 32.0 : dq < -32.0 ? -32.0 : dq;
            const double seed = (double)q0 + dq;
            q0 = seed > 0.0 ? (uint32_t)(seed + 0.5) : 0u;
        }
    }
    q0 = clampu(q0, qp_min, qp_max);
    double prior = 4.0;
    if (!g_model_off && g_hint.bytes > 0u && g_hint.prev_bytes > 0u && g_hint.prev != g_hint.qp) {
        double sl = 0.0;
        if (hint_pair(&g_hint, &sl)) { prior = sl < 1.0 ? 1.0 : sl > 16.0 ? 16.0 : sl; }
    }

    int have_lo = 0, have_hi = 0;
    uint32_t lo = 0, hi = 0, q_final = qp_max;
    int32_t rc = probe(&s, q0);
    if (rc >= 0 && !s.violated) {
        if (s.size <= target) { lo = q0; have_lo = 1; } else { hi = q0; have_hi = 1; }
        for (;;) {
            if (have_lo && (!have_hi ? lo == qp_min : lo == hi + 1u)) { q_final = lo; break; }
            if (have_hi && !have_lo && hi == qp_max) { q_final = qp_max; break; }
            if (s.budget == 0u) { s.violated = 1; break; }
            uint32_t q, bh = 0, bl = 0;
            if (!have_lo) {
                (void)seen(&s, hi, &bh);
                const double dr = target > 0u && bh > target ? log2((double)bh / (double)target) : 16.0;
                const uint32_t dq = clampu((uint32_t)(pair_slope(&s, prior, 1.0, 16.0, 0) * dr + 0.999), 1u, 32u);
                q = hi + dq;
                if (q >= qp_max) { q = qp_max; }
            } else if (!have_hi) {
                (void)seen(&s, lo, &bl);
                const double dr = bl > 0u && (double)target > (double)bl ? log2((double)target / (double)bl) : 16.0;
                const uint32_t dq = clampu((uint32_t)(pair_slope(&s, prior, 1.0, 16.0, 0) * dr + 0.999), 1u, 32u);
                q = dq < lo ? lo - dq : qp_min;
                if (q < qp_min) { q = qp_min; }
                if (q >= lo) { q = lo - 1u; }
            } else {
                (void)seen(&s, hi, &bh);
                (void)seen(&s, lo, &bl);
                double pred;
                if (bh > 0u && bl > 0u && bh > bl) {
                    pred = (double)hi + ((double)lo - (double)hi) *
                                            (log2((double)bh) - log2((double)(target > 0u ? target : 1u))) /
                                            (log2((double)bh) - log2((double)bl));
                } else {
                    pred = (double)((hi + lo) / 2u);
                }
                q = (uint32_t)(pred + 0.5);
                if (q <= hi) { q = hi + 1u; }
                if (q >= lo) { q = lo - 1u; }
                uint32_t tmp = 0;
                if (seen(&s, q, &tmp)) {
                    uint32_t mid = (hi + lo) / 2u;
                    if (mid <= hi || mid >= lo) { mid = hi + 1u; }
                    for (q = mid; q < lo; ++q) {
                        if (!seen(&s, q, &tmp)) { break; }
                    }
                    if (q >= lo) { q_final = lo; break; }
                }
            }
            rc = probe(&s, q);
            if (rc < 0 || s.violated) { break; }
            if (s.size <= target) {
                if (!have_lo || q < lo) { lo = q; }
                have_lo = 1;
            } else {
                if (!have_hi || q > hi) { hi = q; }
                have_hi = 1;
            }
        }
    }
    if (s.violated) { *fallback = 1; return ORC_OK; }
    if (rc < 0) { return rc; }
    rc = emit_final(f, q_final, out, cap, size_out);
    f->count++;
    if (rc < 0) { return rc; }
    *q_used = q_final;
    uint32_t bf = 0;
    if (seen(&s, q_final, &bf)) {
        if (g_hint.valid && g_hint.qp != q_final) {
            g_hint.prev = g_hint.qp;
            g_hint.prev_bytes = g_hint.bytes;
        }
        g_hint.qp = q_final;
        g_hint.valid = 1;
        g_hint.bytes = bf;
    }
    return ORC_OK;
}

static int32_t search_linear(frame_src* f, uint32_t target, uint32_t qp_min, uint32_t qp_max, uint8_t* out,
                             size_t cap, size_t* size_out, uint32_t* q_used)
{
    uint32_t qp = clampu(f->cfg->qp_base, qp_min, qp_max), iters = 0, size = 0;
    int32_t rc = encode_qp(f, qp, &size);
    f->count++;
    iters++;
    while (rc >= 0 && size > target && qp + 8u <= qp_max && iters < 24u) {
        qp += 8u;
        rc = encode_qp(f, qp, &size);
        f->count++;
        iters++;
    }
    while (rc >= 0 && size > target && qp < qp_max && iters < 24u) {
        qp += 1u;
        rc = encode_qp(f, qp, &size);
        f->count++;
        iters++;
    }
    while (rc >= 0 && size <= target && (uint64_t)size * 4u < (uint64_t)target * 3u && qp > qp_min &&
           iters < 24u) {
        const uint32_t cand = qp - 1u;
        const int32_t rc2 = encode_qp(f, cand, &size);
        f->count++;
        iters++;
        if (rc2 >= 0 && size <= target) {
            qp = cand;
            rc = rc2;
            continue;
        }
        rc = encode_qp(f, qp, &size);
        f->count++;
        iters++;
        break;
    }
    *q_used = qp;
    if (rc < 0) { return rc; }
    *size_out = size;
    if (size > cap) { return ORC_EBUF; }
    if (out != NULL) { memcpy(out, f->buf, size); }
    return ORC_OK;
}

static int32_t search_fast(frame_src* f, int exact, uint32_t target, uint32_t qp_min, uint32_t qp_max,
                           uint8_t* out, size_t cap, size_t* size_out, uint32_t* q_used, int* fallback)
{
    probes s;
    memset(&s, 0, sizeof(s));
    s.f = f;
    s.budget = BUDGET;
    const uint32_t q0 = clampu(f->cfg->qp_base, qp_min, qp_max);
    uint32_t q_final = q0;
    int32_t rc = probe(&s, q0);
    if (rc < 0 || s.violated) { goto done; }
    uint32_t qp = q0, iters = 1;
    while (rc >= 0 && s.size > target && qp + 8u <= qp_max && iters < 24u) {
        qp += 8u;
        rc = probe(&s, qp);
        if (rc < 0 || s.violated) { goto done; }
        iters++;
    }
    while (rc >= 0 && s.size > target && qp < qp_max && iters < 24u) {
        qp += 1u;
        rc = probe(&s, qp);
        if (rc < 0 || s.violated) { goto done; }
        iters++;
    }
    const uint32_t q_a = qp;
    q_final = q_a;
    uint32_t extra = 0;
    uint32_t ba = 0;
    (void)seen(&s, q_a, &ba);
    if (ba <= target && (uint64_t)ba * 4u < (uint64_t)target * 3u && q_a > qp_min && iters < 24u) {
        uint32_t lo = q_a, hi = 0;
        int have_notp = 0;
        for (;;) {
            uint32_t next;
            if (!have_notp) {
                double dist = pair_slope(&s, 4.0, 1.0, 16.0, 1) * log2((0.75 * (double)target) / (double)ba);
                if (dist < 0.0) { dist = 0.0; }
                const uint32_t d = (uint32_t)(dist + 0.999);
                next = clampu(lo - d - 1u, qp_min, lo - 1u);
            } else {
                next = (hi + lo) / 2u;
                if (next == hi) { next = hi + 1u; }
                if (next >= lo) { next = lo - 1u; }
            }
            rc = probe(&s, next);
            if (rc < 0 || s.violated) { goto done; }
            const uint32_t bn = s.size;
            if ((uint64_t)bn * 4u < (uint64_t)target * 3u) {
                lo = next;
                ba = bn;
                if (lo == qp_min) { break; }
            } else {
                have_notp = 1;
                hi = next;
                if (hi + 1u == lo) { break; }
            }
            if (have_notp && hi + 1u >= lo) { break; }
            if (!have_notp && lo == qp_min) { break; }
        }
        if (have_notp) {
            uint32_t bc = 0;
            (void)seen(&s, hi, &bc);
            if (bc <= target) {
                q_final = hi;
            } else {
                q_final = hi + 1u;
                extra = 2;
            }
        } else {
            q_final = qp_min;
        }
    }
    if (iters + (q_final < q_a ? q_a - q_final : 0u) + extra > 24u) {
        s.violated = 1;
        goto done;
    }
    if (exact) {
        rc = emit_final(f, q_final, out, cap, size_out);
        f->count++;
    } else {
        if (!s.has_last || s.last_qp != q_final) {
            rc = probe(&s, q_final);
            if (rc < 0 || s.violated) { goto done; }
        }
        uint32_t b = 0;
        (void)seen(&s, q_final, &b);
        *size_out = b;
        if (b > cap) {
            rc = ORC_EBUF;
        } else {
            uint32_t tmp = 0;
            rc = encode_qp(f, q_final, &tmp);   /* the scratch buffer holds the last probe's packet */
            if (rc >= 0 && out != NULL) { memcpy(out, f->buf, b); }
        }
    }
done:
    if (s.violated) { *fallback = 1; return ORC_OK; }
    *q_used = q_final;
    return rc;
}

int32_t orc_frame_encode_sized(const orc_cfg* cfg, const uint16_t* const planes[4], const size_t strides[4],
                               uint32_t target, uint8_t qp_min, uint8_t qp_max, uint8_t* qp_used,
                               uint8_t* out, size_t cap, size_t* size_out, uint32_t* count)
{
    if (g_mode < 0) {
        const char* env = getenv("TOPOS_SIZED_SEARCH");
        g_mode = env == NULL ? 0 : strcmp(env, "linear") == 0 ? 1 : strcmp(env, "r6") == 0 ? 2 : 0;
    }
    const uint32_t em = cfg->reserved[0] & 0xFFu;
    if (em == 9u || em == 10u || em == 11u || cfg->pixel_format == 3u) { return ORC_ENOTIMPL; }
    if (qp_min > QP_MAX || qp_max > QP_MAX || qp_min > qp_max) { return ORC_EINVAL; }
    frame_src f;
    memset(&f, 0, sizeof(f));
    f.cfg = cfg;
    f.planes = planes;
    f.strides = strides;
    const uint64_t blocks = orc_color_blocks(cfg);
    f.cap = (size_t)(blocks * 2048u + 65536u +
                     (cfg->alpha_mode ? (uint64_t)cfg->visible_width * cfg->visible_height * 16u : 0u));
    f.buf = (uint8_t*)malloc(f.cap);
    if (f.buf == NULL) { return ORC_ENOMEM; }
    const int exact = blocks > 0u && blocks * 64u * 4u <= FCACHE_MAX;
    uint32_t q = 0;
    size_t size = 0;
    int fallback = 0;
    int32_t rc;
    if (g_mode == 1) {
        rc = search_linear(&f, target, qp_min, qp_max, out, cap, &size, &q);
    } else {
        rc = g_mode == 2 ? search_fast(&f, exact, target, qp_min, qp_max, out, cap, &size, &q, &fallback)
                         : search_strict(&f, exact, target, qp_min, qp_max, out, cap, &size, &q, &fallback);
        if (fallback) { rc = search_linear(&f, target, qp_min, qp_max, out, cap, &size, &q); }
    }
    free(f.buf);
    if (qp_used != NULL) { *qp_used = (uint8_t)q; }
    if (size_out != NULL) { *size_out = size; }
    if (count != NULL) { *count = f.count; }
    return rc;
}
