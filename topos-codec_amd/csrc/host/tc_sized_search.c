/* The qp search of tc_frame_encode_sized replayed over a table of measured
 * byte counts (reference native/src/codec/codec.c):
 *   :7578-7787  strict (default): seeded start, log-slope jumps, bracket
 *               interpolation; the smallest qp whose packet fits
 *   :7554-7576  strict_slope_from_pts; :7375-7395 sized_fit_slope
 *   :7199-7265  legacy linear (+8, +1, reclaim while below 3/4 of target)
 *   :7404-7551  r6 fast search
 *   :7303-7351  probe bookkeeping: 24-probe budget, monotonicity violations
 *   :7765-7785  seed state: the last two working points (qp, bytes)
 * Pure host code: a probe of an unmeasured qp stops the replay with SZ_MISS;
 * the caller measures it and replays from the start. */
#include <math.h>
#include <string.h>

#include "tc_sized.h"



typedef struct sz_probes {
    uint32_t qp[SZ_BUDGET + 1], b[SZ_BUDGET + 1], n;
    uint32_t budget, last_qp, size;
    int has_last, violated;
} sz_probes;

static int sz_known(sz_item* it, uint32_t q)
{
    if (!it->known[q]) { it->miss = q; return 0; }
    return 1;
}

/* a search probe (codec.c:7303-7351): budget, bookkeeping, violations */
static int32_t sz_probe(sz_item* it, sz_probes* s, uint32_t q)
{
    if (s->budget == 0u) { s->violated = 1; return TC_ERR_LIMIT_EXCEEDED; }
    if (!sz_known(it, q)) { return SZ_MISS; }
    it->count++;
    s->budget--;
    s->last_qp = q;
    s->has_last = 1;
    if (it->rc[q] < 0) { return it->rc[q]; }
    const uint32_t b = it->bytes[q];
    s->size = b;
    for (uint32_t i = 0; i < s->n; ++i) {
        if ((s->qp[i] < q && s->b[i] < b) || (s->qp[i] > q && s->b[i] > b)) { s->violated = 1; }
    }
    s->qp[s->n] = q;
    s->b[s->n] = b;
    s->n++;
    return TC_OK;
}

static int sz_seen(const sz_probes* s, uint32_t q, uint32_t* b)
{
    for (uint32_t i = 0; i < s->n; ++i) {
        if (s->qp[i] == q) { *b = s->b[i]; return 1; }
    }
    return 0;
}

/* newest probe pair with qp up and bytes down by more than 2^0.05:
 * strict_slope_from_pts (need_b0 = 0) and sized_fit_slope (need_b0 = 1) */
static double sz_slope(const sz_probes* s, double dflt, int need_b0)
{
    for (uint32_t i = s->n; i-- > 1u;) {
        for (uint32_t j = i; j-- > 0u;) {
            const uint32_t q1 = s->qp[i], b1 = s->b[i], q0 = s->qp[j], b0 = s->b[j];
            if (q1 > q0 && b0 > b1 && (need_b0 ? b0 > 0u : b1 > 0u)) {
                const double dr = log2((double)b0 / (double)b1);
                if (dr > 0.05) {
                    const double v = (double)(q1 - q0) / dr;
                    return v < 1.0 ? 1.0 : v > 16.0 ? 16.0 : v;
                }
            }
        }
    }
    return dflt;
}

static uint32_t sz_clamp(uint32_t v, uint32_t lo, uint32_t hi) { return v < lo ? lo : v > hi ? hi : v; }

/* slope through the seed's two working points, when they make one */
static int sz_seed_slope(const tcx_sized_stream* h, double* out)
{
    const uint32_t q_hi = h->qp > h->prev ? h->qp : h->prev;
    const uint32_t q_lo = q_hi == h->qp ? h->prev : h->qp;
    const uint32_t b_hi = q_hi == h->qp ? h->bytes : h->prev_bytes;
    const uint32_t b_lo = q_hi == h->qp ? h->prev_bytes : h->bytes;
    if (q_hi > q_lo && b_lo > b_hi && b_hi > 0u) {
        *out = (double)(q_hi - q_lo) / log2((double)b_lo / (double)b_hi);
        return 1;
    }
    return 0;
}

/* the final packet at q: its size decides BUFFER_TOO_SMALL */
static int32_t sz_final(sz_item* it, uint32_t q)
{
    if (!sz_known(it, q)) { return SZ_MISS; }
    it->count++;
    if (it->rc[q] < 0) { return it->rc[q]; }
    it->q_fin = q;
    it->size = it->bytes[q];
    return it->bytes[q] > it->out_cap ? TC_ERR_BUFFER_TOO_SMALL : TC_OK;
}

#define TRY(expr)                                  \
    do {                                           \
        rc = (expr);                               \
        if (rc == SZ_MISS) { return SZ_MISS; }     \
    } while (0)

static int32_t sz_strict(sz_item* it, int* fallback)
{
    if (!it->exact) { *fallback = 1; return TC_OK; }
    sz_probes s;
    memset(&s, 0, sizeof(s));
    s.budget = SZ_BUDGET;
    const tcx_sized_stream* h = &it->seed;
    uint32_t q0 = it->cfg.qp_base;
    if (it->seed_on && h->valid) {
        q0 = h->qp;
        if (it->model_on && h->bytes > 0u && it->target > 0u) {
            const double dr = log2((double)h->bytes / (double)it->target);
            double slope = 0.0;
            if (h->prev_bytes > 0u && h->prev != h->qp) { (void)sz_seed_slope(h, &slope); }
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
    q0 = sz_clamp(q0, it->qp_min, it->qp_max);
    double prior = 4.0;
    if (it->model_on && h->bytes > 0u && h->prev_bytes > 0u && h->prev != h->qp) {
        double sl = 0.0;
        if (sz_seed_slope(h, &sl)) { prior = sl < 1.0 ? 1.0 : sl > 16.0 ? 16.0 : sl; }
    }
    const uint32_t target = it->target, qp_min = it->qp_min, qp_max = it->qp_max;
    int have_lo = 0, have_hi = 0;
    uint32_t lo = 0, hi = 0, q_final = qp_max;
    int32_t rc;
    TRY(sz_probe(it, &s, q0));
    if (rc >= 0 && !s.violated) {
        if (s.size <= target) { lo = q0; have_lo = 1; } else { hi = q0; have_hi = 1; }
        for (;;) {
            if (have_lo && (!have_hi ? lo == qp_min : lo == hi + 1u)) { q_final = lo; break; }
            if (have_hi && !have_lo && hi == qp_max) { q_final = qp_max; break; }
            if (s.budget == 0u) { s.violated = 1; break; }
            uint32_t q, bh = 0, bl = 0;
            if (!have_lo) {
                (void)sz_seen(&s, hi, &bh);
                const double dr = target > 0u && bh > target ? log2((double)bh / (double)target) : 16.0;
                q = hi + sz_clamp((uint32_t)(sz_slope(&s, prior, 0) * dr + 0.999), 1u, 32u);
                if (q >= qp_max) { q = qp_max; }
            } else if (!have_hi) {
                (void)sz_seen(&s, lo, &bl);
                const double dr = bl > 0u && (double)target > (double)bl ? log2((double)target / (double)bl) : 16.0;
                const uint32_t dq = sz_clamp((uint32_t)(sz_slope(&s, prior, 0) * dr + 0.999), 1u, 32u);
                q = dq < lo ? lo - dq : qp_min;
                if (q < qp_min) { q = qp_min; }
                if (q >= lo) { q = lo - 1u; }
            } else {
                (void)sz_seen(&s, hi, &bh);
                (void)sz_seen(&s, lo, &bl);
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
                if (sz_seen(&s, q, &tmp)) {
                    uint32_t mid = (hi + lo) / 2u;
                    if (mid <= hi || mid >= lo) { mid = hi + 1u; }
                    for (q = mid; q < lo; ++q) {
                        if (!sz_seen(&s, q, &tmp)) { break; }
                    }
                    if (q >= lo) { q_final = lo; break; }
                }
            }
            TRY(sz_probe(it, &s, q));
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
    if (s.violated) { *fallback = 1; return TC_OK; }
    if (rc < 0) { return rc; }
    TRY(sz_final(it, q_final));
    if (rc < 0) { return rc; }
    it->q_used = q_final;
    uint32_t bf = 0;
    if (sz_seen(&s, q_final, &bf)) {
        it->seed_out = it->seed;
        if (it->seed_out.valid && it->seed_out.qp != q_final) {
            it->seed_out.prev = it->seed_out.qp;
            it->seed_out.prev_bytes = it->seed_out.bytes;
        }
        it->seed_out.qp = q_final;
        it->seed_out.valid = 1;
        it->seed_out.bytes = bf;
        it->seed_update = 1;
    }
    return TC_OK;
}

/* legacy linear probe: a whole-frame encode's byte count */
static int32_t sz_lin(sz_item* it, uint32_t q, uint32_t* size)
{
    if (!sz_known(it, q)) { return SZ_MISS; }
    it->count++;
    if (it->rc[q] < 0) { return it->rc[q]; }
    *size = it->bytes[q];
    return TC_OK;
}

static int32_t sz_linear(sz_item* it)
{
    const uint32_t target = it->target, qp_min = it->qp_min, qp_max = it->qp_max;
    uint32_t qp = sz_clamp(it->cfg.qp_base, qp_min, qp_max), iters = 0, size = 0;
    int32_t rc;
    TRY(sz_lin(it, qp, &size));
    iters++;
    while (rc >= 0 && size > target && qp + 8u <= qp_max && iters < 24u) {
        qp += 8u;
        TRY(sz_lin(it, qp, &size));
        iters++;
    }
    while (rc >= 0 && size > target && qp < qp_max && iters < 24u) {
        qp += 1u;
        TRY(sz_lin(it, qp, &size));
        iters++;
    }
    while (rc >= 0 && size <= target && (uint64_t)size * 4u < (uint64_t)target * 3u && qp > qp_min &&
           iters < 24u) {
        const uint32_t cand = qp - 1u;
        int32_t rc2;
        rc2 = sz_lin(it, cand, &size);
        if (rc2 == SZ_MISS) { return SZ_MISS; }
        iters++;
        if (rc2 >= 0 && size <= target) {
            qp = cand;
            rc = rc2;
            continue;
        }
        TRY(sz_lin(it, qp, &size));
        iters++;
        break;
    }
    it->q_used = qp;
    if (rc < 0) { return rc; }
    it->q_fin = qp;
    it->size = size;
    return size > it->out_cap ? TC_ERR_BUFFER_TOO_SMALL : TC_OK;
}

static int32_t sz_fast(sz_item* it, int* fallback)
{
    sz_probes s;
    memset(&s, 0, sizeof(s));
    s.budget = SZ_BUDGET;
    const uint32_t target = it->target, qp_min = it->qp_min, qp_max = it->qp_max;
    const uint32_t q0 = sz_clamp(it->cfg.qp_base, qp_min, qp_max);
    uint32_t q_final = q0, qp = q0, iters = 1, q_a = q0, extra = 0, ba = 0;
    int32_t rc;
    TRY(sz_probe(it, &s, q0));
    if (rc < 0 || s.violated) { goto done; }
    while (rc >= 0 && s.size > target && qp + 8u <= qp_max && iters < 24u) {
        qp += 8u;
        TRY(sz_probe(it, &s, qp));
        if (rc < 0 || s.violated) { goto done; }
        iters++;
    }
    while (rc >= 0 && s.size > target && qp < qp_max && iters < 24u) {
        qp += 1u;
        TRY(sz_probe(it, &s, qp));
        if (rc < 0 || s.violated) { goto done; }
        iters++;
    }
    q_a = qp;
    q_final = q_a;
    (void)sz_seen(&s, q_a, &ba);
    if (ba <= target && (uint64_t)ba * 4u < (uint64_t)target * 3u && q_a > qp_min && iters < 24u) {
        uint32_t lo = q_a, hi = 0;
        int have_notp = 0;
        for (;;) {
            uint32_t next;
            if (!have_notp) {
                double dist = sz_slope(&s, 4.0, 1) * log2((0.75 * (double)target) / (double)ba);
                if (dist < 0.0) { dist = 0.0; }
                const uint32_t d = (uint32_t)(dist + 0.999);
                next = sz_clamp(lo - d - 1u, qp_min, lo - 1u);
            } else {
                next = (hi + lo) / 2u;
                if (next == hi) { next = hi + 1u; }
                if (next >= lo) { next = lo - 1u; }
            }
            TRY(sz_probe(it, &s, next));
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
            (void)sz_seen(&s, hi, &bc);
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
    if (it->exact) {
        TRY(sz_final(it, q_final));
    } else {
        if (!s.has_last || s.last_qp != q_final) {
            TRY(sz_probe(it, &s, q_final));
            if (rc < 0 || s.violated) { goto done; }
        }
        it->q_fin = q_final;
        it->size = it->bytes[q_final];
        rc = it->size > it->out_cap ? TC_ERR_BUFFER_TOO_SMALL : TC_OK;
    }
done:
    if (s.violated) { *fallback = 1; return TC_OK; }
    it->q_used = q_final;
    return rc;
}

/* replays the whole search; SZ_MISS leaves the needed qp in it->miss */
int32_t sz_replay(sz_item* it, int mode)
{
    it->count = 0;
    it->q_used = 0;
    it->q_fin = 0;
    it->size = 0;
    it->seed_update = 0;
    if (mode == 1) { return sz_linear(it); }
    int fallback = 0;
    const int32_t rc = mode == 2 ? sz_fast(it, &fallback) : sz_strict(it, &fallback);
    if (rc == SZ_MISS) { return SZ_MISS; }
    if (fallback) {
        it->seed_update = 0;
        return sz_linear(it);
    }
    return rc;
}


/* candidate qps for a missing item: the needed one and its nearest
 * unmeasured neighbours inside [qp_min, qp_max] */
uint32_t sz_candidates(const sz_item* it, uint32_t want, uint32_t* out)
{
    uint32_t n = 0;
    out[n++] = it->miss;
    for (uint32_t d = 1; n < want && d <= SZ_QP_MAX; ++d) {
        const uint32_t up = it->miss + d;
        if (up <= it->qp_max && !it->known[up]) { out[n++] = up; }
        if (n < want && it->miss >= it->qp_min + d && !it->known[it->miss - d]) { out[n++] = it->miss - d; }
    }
    return n;
}

void sz_item_init(sz_item* it, const topos_frame_config* cfg, uint32_t target, uint32_t qp_min, uint32_t qp_max,
                  size_t out_cap, int exact, const tcx_sized_stream* seed, int seed_on, int model_on)
{
    memset(it, 0, sizeof(*it));
    it->cfg = *cfg;
    it->target = target;
    it->qp_min = qp_min;
    it->qp_max = qp_max;
    it->out_cap = out_cap;
    it->exact = exact;
    if (seed != NULL) { it->seed = *seed; }
    it->seed_on = seed != NULL && seed_on;
    it->model_on = model_on;
}
