/* Serving loop (include/topos_gpu.h: tcx_server_*): continuous batching of
 * encode and decode requests over the device engine.
 *
 * One scheduler thread owns the engine lanes. Each turn it
 *  1. admits waiting requests in arrival order while the HBM budget (one
 *     device allocation managed by tc_arena) can hold their footprint:
 *     encode = worst-case packet + staged host inputs, decode = output
 *     planes + staged packet;
 *  2. starts a launch sequence on every free lane: the direction whose oldest
 *     admitted request arrived first, its admitted requests in arrival order,
 *     mixed sizes and profiles in one batch. A queue longer than max_batch is
 *     cut into equal batches (160 waiting with max_batch 96: 80 and 80), so
 *     the last one is not a remainder that holds a lane for a whole serial
 *     entropy chain with a fraction of the device busy. While requests keep
 *     arriving less than gather_us apart, the launch waits for them (at most
 *     20 x gather_us after its oldest request arrived): a burst then forms
 *     equal full batches instead of a first batch of one frame, and isolated
 *     arrivals are not delayed;
 *  3. polls the running lanes without blocking; a finished lane's requests
 *     complete, their reservations shrink to what the result holds (packet
 *     bytes / output planes) and the staged inputs return to the budget.
 * Submission, polling and release take the server lock only briefly; engine
 * calls run outside it. Lane contexts are taken with the non-blocking engine
 * acquire, so the scheduler never waits on a context it holds itself.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "tc_arena.h"
#include "tc_host.h"
#include "topos_gpu.h"

#define SRV_MAX_LANES 4u
#define SRV_MAX_BATCH 256u
#define SRV_SLACK 1024u /* tail slack of a packet region: the decoder reads through aligned windows */

enum { RQ_WAIT = 0, RQ_ADMITTED, RQ_RUNNING, RQ_DONE, RQ_POLLED };

typedef struct sreq {
    uint64_t id;
    tcx_request r;
    tc_view* view;              /* decode: parsed structure */
    int state;
    size_t off, len;            /* arena range (len 0: nothing held) */
    size_t keep;                /* bytes of the range kept once the request completes */
    uint64_t in[4];             /* encode: staged input planes */
    size_t in_pitch[4];         /* their pitch (elements) */
    uint64_t out_pkt;           /* encode: packet buffer */
    size_t out_cap;
    uint64_t pkt;               /* decode: the packet in HBM */
    uint64_t planes[4];         /* decode: output planes */
    size_t pstride[4];
    uint32_t nplanes;
    int32_t status;
    uint32_t packet_size, batch_frames;
    int64_t t_submit, t_admit, t_launch, t_done;
    topos_frame_output info;    /* decode, when the caller passed none */
    topos_frame_stats stats;    /* encode, when the caller passed none */
    struct sreq* next;          /* queue link */
    struct sreq* live_next;     /* registry of requests not yet released */
} sreq;

typedef struct squeue { sreq* head; sreq* tail; uint32_t n; } squeue;

typedef struct lane {
    int busy;
    uint32_t kind, n;
    tc_ectx* c;
    sreq* items[SRV_MAX_BATCH];
    tc_enc_req er[SRV_MAX_BATCH];
    tc_dec_req dr[SRV_MAX_BATCH];
    int32_t rc;
} lane;

struct tcx_server {
    pthread_mutex_t mu;
    pthread_cond_t cv_work;     /* scheduler: submissions, releases, stop */
    pthread_cond_t cv_done;     /* pollers: completions */
    pthread_t thr;
    int started, stop;
    uint32_t max_batch, max_inflight;
    int64_t gather_ns;          /* arrival coalescing gap (0: off) */
    int64_t last_submit;        /* time of the latest submission */
    int64_t wake;               /* a deferred launch is due by then (0: none) */
    uint64_t base;              /* the budget's device allocation */
    tc_arena arena;
    squeue q_wait, q_enc, q_dec, q_done;
    sreq* live;
    uint64_t next_id;
    uint32_t nbusy;
    lane lanes[SRV_MAX_LANES];
    tcx_server_stats st;
};

static size_t al256(size_t v) { return (v + 255u) & ~(size_t)255u; }

int64_t tcx_now_ns(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (int64_t)ts.tv_sec * 1000000000LL + ts.tv_nsec;
}

static void q_push(squeue* q, sreq* r)
{
    r->next = NULL;
    if (q->tail) { q->tail->next = r; } else { q->head = r; }
    q->tail = r;
    q->n++;
}

static sreq* q_pop(squeue* q)
{
    sreq* r = q->head;
    if (r) {
        q->head = r->next;
        if (q->head == NULL) { q->tail = NULL; }
        r->next = NULL;
        q->n--;
    }
    return r;
}

static void req_free(sreq* q)
{
    free(q->view);
    free(q);
}

/* ------------------------------------------------------------ admission */

/* footprint of a request in the budget and the part kept after completion */
static int32_t plan_footprint(sreq* q)
{
    tcx_request* r = &q->r;
    q->len = 0u;
    q->keep = 0u;
    if (r->kind == TCX_REQ_ENCODE) {
        size_t need = 0u;
        if (r->dev_out == 0u) {
            const size_t cap = tc_engine_encode_cap(&r->cfg);
            if (cap == 0u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "encode config has no packet bound"); }
            q->out_cap = cap;
            need = al256(cap + SRV_SLACK);
        } else {
            q->out_cap = r->dev_out_cap;
        }
        if (r->host_planes[0] != NULL) {
            tc_fhdr fh;
            tc_cfg_to_fhdr(&r->cfg, &fh);
            (void)tc_fhdr_geometry(&fh);
            for (uint32_t p = 0; p < fh.plane_count; ++p) {
                q->in_pitch[p] = ((size_t)fh.pvw[p] + 7u) & ~(size_t)7u;
                need += al256(q->in_pitch[p] * fh.pvh[p] * 2u);
            }
        }
        q->len = need;
        return TC_OK;
    }
    const tc_fhdr* fh = &q->view->fh;
    size_t need = 0u;
    q->nplanes = fh->plane_count;
    if (r->dev_out_planes[0] == 0u) {
        for (uint32_t p = 0; p < fh->plane_count; ++p) { need += al256((size_t)fh->pvw[p] * fh->pvh[p] * 2u); }
    }
    q->keep = need;
    if (r->dev_packet == 0u) { need += al256(r->packet_size + SRV_SLACK); }
    q->len = need;
    return TC_OK;
}

/* gives a just-admitted request its device addresses inside the range */
static void place(tcx_server* s, sreq* q)
{
    const uint64_t at = s->base + q->off;
    tcx_request* r = &q->r;
    if (r->kind == TCX_REQ_ENCODE) {
        size_t o = 0u;
        if (r->dev_out == 0u) {
            q->out_pkt = at;
            o = al256(q->out_cap + SRV_SLACK);
        } else {
            q->out_pkt = r->dev_out;
        }
        if (r->host_planes[0] != NULL) {
            tc_fhdr fh;
            tc_cfg_to_fhdr(&r->cfg, &fh);
            (void)tc_fhdr_geometry(&fh);
            for (uint32_t p = 0; p < fh.plane_count; ++p) {
                q->in[p] = at + o;
                o += al256(q->in_pitch[p] * fh.pvh[p] * 2u);
            }
        } else {
            for (uint32_t p = 0; p < 4u; ++p) {
                q->in[p] = r->dev_planes[p];
                q->in_pitch[p] = r->strides[p];
            }
        }
        return;
    }
    const tc_fhdr* fh = &q->view->fh;
    size_t o = 0u;
    for (uint32_t p = 0; p < fh->plane_count; ++p) {
        if (r->dev_out_planes[0] == 0u) {
            q->planes[p] = at + o;
            q->pstride[p] = fh->pvw[p];
            o += al256((size_t)fh->pvw[p] * fh->pvh[p] * 2u);
        } else {
            q->planes[p] = r->dev_out_planes[p];
            q->pstride[p] = r->out_strides[p] ? r->out_strides[p] : fh->pvw[p];
        }
    }
    q->pkt = r->dev_packet ? r->dev_packet : at + o;
}

static int admit_locked(tcx_server* s)
{
    int progress = 0;
    while (s->q_wait.head != NULL) {
        sreq* q = s->q_wait.head;
        if (q->len != 0u && tc_arena_alloc(&s->arena, q->len, &q->off) != 0) { break; }
        (void)q_pop(&s->q_wait);
        q->state = RQ_ADMITTED;
        q->t_admit = tcx_now_ns();
        place(s, q);
        q_push(q->r.kind == TCX_REQ_ENCODE ? &s->q_enc : &s->q_dec, q);
        progress = 1;
    }
    return progress;
}

/* ------------------------------------------------------------ launches */

static int32_t lane_launch(lane* L)
{
    tcg_ctx* g = tc_engine_stream(L->c);
    if (L->kind == TCX_REQ_ENCODE) {
        for (uint32_t k = 0; k < L->n; ++k) {
            sreq* q = L->items[k];
            tc_enc_req* e = &L->er[k];
            memset(e, 0, sizeof(*e));
            if (q->r.host_planes[0] != NULL) {
                tc_fhdr fh;
                tc_cfg_to_fhdr(&q->r.cfg, &fh);
                (void)tc_fhdr_geometry(&fh);
                for (uint32_t p = 0; p < fh.plane_count; ++p) {
                    const size_t vw = fh.pvw[p], sst = q->r.strides[p] ? q->r.strides[p] : vw;
                    if (tcg_h2d_2d(g, q->in[p], q->in_pitch[p] * 2u, q->r.host_planes[p], sst * 2u, vw * 2u,
                                   fh.pvh[p]) != 0) {
                        return tc_fail(TC_ERR_OUT_OF_MEMORY, "input staging: %s", tcg_last_error());
                    }
                }
            }
            e->cfg = &q->r.cfg;
            e->dev_planes = q->in;
            e->dev_strides = q->in_pitch;
            e->dev_out = q->out_pkt;
            e->dev_out_cap = q->out_cap;
            e->out_cap = q->out_cap;
            e->stats = q->r.stats ? q->r.stats : &q->stats;
        }
        return tc_engine_encode_submit(L->c, L->er, L->n);
    }
    for (uint32_t k = 0; k < L->n; ++k) {
        sreq* q = L->items[k];
        tc_dec_req* d = &L->dr[k];
        memset(d, 0, sizeof(*d));
        if (q->r.dev_packet == 0u && tcg_h2d(g, q->pkt, q->r.packet, q->r.packet_size) != 0) {
            return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet staging: %s", tcg_last_error());
        }
        d->data = q->r.packet;
        d->size = q->r.packet_size;
        d->dev_data = q->pkt;
        d->dev_planes = q->planes;
        d->dev_strides = q->pstride;
        d->info = q->r.info ? q->r.info : &q->info;
        d->view = q->view;
    }
    return tc_engine_decode_submit(L->c, L->dr, L->n);
}

/* runs on the scheduler thread without the server lock; the lane's requests
 * are owned by the scheduler while they run */
static void lane_finish(lane* L)
{
    if (L->kind == TCX_REQ_ENCODE) {
        (void)tc_engine_encode_finish(L->c, L->er, L->n);
        for (uint32_t k = 0; k < L->n; ++k) {
            L->items[k]->status = L->er[k].rc;
            L->items[k]->packet_size = L->er[k].packet_size;
        }
    } else {
        (void)tc_engine_decode_finish(L->c, L->dr, L->n);
        for (uint32_t k = 0; k < L->n; ++k) { L->items[k]->status = L->dr[k].rc; }
    }
    tc_engine_release(L->c);
    L->c = NULL;
}

static void complete_locked(tcx_server* s, sreq* q, int32_t status)
{
    if (status != 0) { q->status = status; }
    q->t_done = tcx_now_ns();
    if (q->len != 0u) {
        size_t keep = q->keep;
        if (q->r.kind == TCX_REQ_ENCODE && q->r.dev_out == 0u) {
            keep = q->status >= TC_OK ? al256((size_t)q->packet_size + SRV_SLACK) : 0u;
        }
        if (keep == 0u) {
            (void)tc_arena_free(&s->arena, q->off, q->len);
            q->len = 0u;
        } else if (keep < q->len) {
            (void)tc_arena_trim(&s->arena, q->off, q->len, keep);
            q->len = keep;
        }
    }
    q->state = RQ_DONE;
    q_push(&s->q_done, q);
    s->st.completed++;
}

/* starts launch sequences on free lanes; called with the lock held, drops it
 * around engine calls */
static int launch_locked(tcx_server* s)
{
    int progress = 0;
    while (s->nbusy < s->max_inflight && (s->q_enc.head != NULL || s->q_dec.head != NULL)) {
        lane* L = NULL;
        for (uint32_t i = 0; i < s->max_inflight; ++i) {
            if (!s->lanes[i].busy) { L = &s->lanes[i]; break; }
        }
        if (L == NULL) { break; }
        tc_ectx* c = NULL;
        const int32_t arc = tc_engine_try_acquire(&c);
        if (arc == 1) { break; } /* every engine context is taken: wait for a lane to finish */
        squeue* src;
        if (s->q_enc.head == NULL) { src = &s->q_dec; }
        else if (s->q_dec.head == NULL) { src = &s->q_enc; }
        else { src = s->q_enc.head->t_submit <= s->q_dec.head->t_submit ? &s->q_enc : &s->q_dec; }
        const int64_t now = tcx_now_ns();
        if (s->gather_ns > 0 && !s->stop && now - s->last_submit < s->gather_ns &&
            now - src->head->t_submit < 20 * s->gather_ns) {
            /* arrivals are still streaming in: wait for them */
            const int64_t due_gap = s->last_submit + s->gather_ns, due_max = src->head->t_submit + 20 * s->gather_ns;
            s->wake = due_gap < due_max ? due_gap : due_max;
            if (arc == TC_OK) { tc_engine_release(c); }
            break;
        }
        s->wake = 0;
        /* equal batches: ceil(n / ceil(n / max_batch)) */
        const uint32_t nbat = (src->n + s->max_batch - 1u) / s->max_batch;
        const uint32_t take = (src->n + nbat - 1u) / nbat;
        L->kind = src == &s->q_enc ? TCX_REQ_ENCODE : TCX_REQ_DECODE;
        L->n = 0;
        while (L->n < take && src->head != NULL) {
            sreq* q = q_pop(src);
            q->state = RQ_RUNNING;
            q->t_launch = now;
            L->items[L->n++] = q;
        }
        for (uint32_t k = 0; k < L->n; ++k) { L->items[k]->batch_frames = L->n; }
        L->busy = 1;
        s->nbusy++;
        s->st.batches++;
        s->st.batch_frames += L->n;
        progress = 1;
        if (arc != TC_OK) {
            for (uint32_t k = 0; k < L->n; ++k) { complete_locked(s, L->items[k], arc); }
            L->busy = 0;
            s->nbusy--;
            pthread_cond_broadcast(&s->cv_done);
            continue;
        }
        L->c = c;
        pthread_mutex_unlock(&s->mu);
        const int32_t rc = lane_launch(L);
        if (rc != TC_OK) {
            (void)tcg_sync(tc_engine_stream(L->c));
            tc_engine_release(L->c);
            L->c = NULL;
        }
        pthread_mutex_lock(&s->mu);
        if (rc != TC_OK) {
            for (uint32_t k = 0; k < L->n; ++k) { complete_locked(s, L->items[k], rc); }
            L->busy = 0;
            s->nbusy--;
            pthread_cond_broadcast(&s->cv_done);
        }
    }
    return progress;
}

static void* srv_main(void* arg)
{
    tcx_server* s = (tcx_server*)arg;
    pthread_mutex_lock(&s->mu);
    for (;;) {
        int progress = admit_locked(s);
        progress |= launch_locked(s);
        if (s->stop && s->nbusy == 0u && s->q_enc.head == NULL && s->q_dec.head == NULL) {
            /* what the budget cannot take now never fits: nothing is released
             * while the server shuts down */
            while (s->q_wait.head != NULL) {
                sreq* q = q_pop(&s->q_wait);
                q->len = 0u;
                complete_locked(s, q, TC_ERR_STATE);
            }
            pthread_cond_broadcast(&s->cv_done);
            break;
        }
        if (s->nbusy > 0u) {
            lane* fin[SRV_MAX_LANES];
            uint32_t nf = 0;
            pthread_mutex_unlock(&s->mu);
            for (uint32_t i = 0; i < s->max_inflight; ++i) {
                lane* L = &s->lanes[i];
                if (L->busy && tcg_query(tc_engine_stream(L->c)) != 1) {
                    lane_finish(L);
                    fin[nf++] = L;
                }
            }
            pthread_mutex_lock(&s->mu);
            for (uint32_t j = 0; j < nf; ++j) {
                lane* L = fin[j];
                for (uint32_t k = 0; k < L->n; ++k) { complete_locked(s, L->items[k], 0); }
                L->busy = 0;
                s->nbusy--;
            }
            if (nf) {
                pthread_cond_broadcast(&s->cv_done);
                progress = 1;
            }
        }
        if (!progress) {
            if (s->nbusy > 0u || s->wake != 0) {
                /* lanes running or a launch deferred: look again shortly */
                int64_t ns = 20000;
                if (s->nbusy == 0u) {
                    const int64_t left = s->wake - tcx_now_ns();
                    ns = left < 1000 ? 1000 : (left < 50000 ? left : 50000);
                }
                pthread_mutex_unlock(&s->mu);
                const struct timespec ts = { 0, (long)ns };
                nanosleep(&ts, NULL);
                pthread_mutex_lock(&s->mu);
            } else {
                pthread_cond_wait(&s->cv_work, &s->mu);
            }
        }
    }
    pthread_mutex_unlock(&s->mu);
    return NULL;
}

/* ------------------------------------------------------------ public API */

int32_t tcx_server_create(const tcx_server_config* cfg, tcx_server** out)
{
    if (cfg == NULL || out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "server config or out NULL"); }
    *out = NULL;
    const uint32_t mb = cfg->max_batch ? cfg->max_batch : 8u;
    const uint32_t mi = cfg->max_inflight ? cfg->max_inflight : SRV_MAX_LANES;
    if (mb > SRV_MAX_BATCH || mi > SRV_MAX_LANES) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "max_batch %u (1..%u) or max_inflight %u (1..%u) out of range", mb,
                       SRV_MAX_BATCH, mi, SRV_MAX_LANES);
    }
    if (cfg->hbm_budget < 4096u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "hbm_budget %zu too small", cfg->hbm_budget); }
    tc_ectx* c = NULL;
    int32_t rc = tc_engine_acquire(&c); /* initialises the device (fails loudly without one) */
    if (rc != TC_OK) { return rc; }
    tc_engine_release(c);
    tcx_server* s = (tcx_server*)calloc(1, sizeof(tcx_server));
    if (s == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "server"); }
    s->max_batch = mb;
    s->max_inflight = mi;
    s->gather_ns = (int64_t)cfg->gather_us * 1000;
    if (tc_arena_init(&s->arena, cfg->hbm_budget) != 0) { free(s); return tc_fail(TC_ERR_OUT_OF_MEMORY, "arena"); }
    if (tcg_malloc(&s->base, s->arena.size + SRV_SLACK) != 0) {
        tc_arena_fini(&s->arena);
        free(s);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "HBM budget of %zu bytes: %s", cfg->hbm_budget, tcg_last_error());
    }
    pthread_mutex_init(&s->mu, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    s->next_id = 1u;
    s->st.budget_bytes = s->arena.size;
    if (pthread_create(&s->thr, NULL, srv_main, s) != 0) {
        (void)tcg_free(s->base);
        tc_arena_fini(&s->arena);
        free(s);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "scheduler thread");
    }
    s->started = 1;
    *out = s;
    return TC_OK;
}

int32_t tcx_server_submit(tcx_server* s, const tcx_request* req, uint64_t* id)
{
    if (s == NULL || req == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "server or request NULL"); }
    sreq* q = (sreq*)calloc(1, sizeof(sreq));
    if (q == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "request"); }
    q->r = *req;
    q->t_submit = tcx_now_ns();
    int32_t rc = TC_OK;
    if (req->kind == TCX_REQ_ENCODE) {
        extern int32_t tc_enc_check_dev(const topos_frame_config*, const uint64_t*, const size_t*);
        if (req->host_planes[0] != NULL) {
            for (uint32_t p = 0; p < 4u; ++p) { q->r.dev_planes[p] = (uint64_t)(uintptr_t)req->host_planes[p]; }
        }
        rc = tc_enc_check_dev(&q->r.cfg, q->r.dev_planes, q->r.strides);
        if (req->host_planes[0] != NULL) { memset(q->r.dev_planes, 0, sizeof(q->r.dev_planes)); }
        if (rc == TC_OK && req->dev_out != 0u && req->dev_out_cap == 0u) {
            rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "dev_out given with capacity 0");
        }
    } else if (req->kind == TCX_REQ_DECODE) {
        q->view = (tc_view*)malloc(sizeof(tc_view));
        if (q->view == NULL) {
            rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view");
        } else {
            rc = tc_dec_prepare_check(req->packet, req->packet_size, req->info ? req->info : &q->info, q->view);
        }
        if (rc == TC_OK && req->dev_out_planes[0] != 0u) {
            for (uint32_t p = 0; p < q->view->fh.plane_count; ++p) {
                if (req->dev_out_planes[p] == 0u) {
                    rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "dev_out_planes[%u] == 0", p);
                } else if (req->out_strides[p] != 0u && req->out_strides[p] < q->view->fh.pvw[p]) {
                    rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "out_strides[%u] below the visible width", p);
                }
            }
        }
    } else {
        rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "request kind %u", req->kind);
    }
    if (rc == TC_OK) { rc = plan_footprint(q); }
    if (rc == TC_OK && q->len > s->arena.size) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "request needs %zu bytes of HBM, the budget is %zu", q->len,
                     s->arena.size);
    }
    if (rc != TC_OK) { req_free(q); return rc; }
    pthread_mutex_lock(&s->mu);
    if (s->stop) {
        pthread_mutex_unlock(&s->mu);
        req_free(q);
        return tc_fail(TC_ERR_STATE, "server is stopping");
    }
    q->id = s->next_id++;
    q->live_next = s->live;
    s->live = q;
    q_push(&s->q_wait, q);
    s->last_submit = tcx_now_ns();
    s->st.submitted++;
    if (id) { *id = q->id; }
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    return TC_OK;
}

int32_t tcx_server_poll(tcx_server* s, tcx_completion* out, uint32_t max, uint32_t* n, int64_t timeout_ns)
{
    if (s == NULL || n == NULL || (out == NULL && max > 0u)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "poll arguments NULL");
    }
    *n = 0;
    pthread_mutex_lock(&s->mu);
    if (s->q_done.head == NULL && timeout_ns != 0) {
        struct timespec dl;
        clock_gettime(CLOCK_REALTIME, &dl);
        if (timeout_ns > 0) {
            const int64_t t = (int64_t)dl.tv_nsec + timeout_ns;
            dl.tv_sec += (time_t)(t / 1000000000LL);
            dl.tv_nsec = (long)(t % 1000000000LL);
        }
        while (s->q_done.head == NULL) {
            if (timeout_ns < 0) {
                pthread_cond_wait(&s->cv_done, &s->mu);
            } else if (pthread_cond_timedwait(&s->cv_done, &s->mu, &dl) != 0) {
                break;
            }
        }
    }
    while (*n < max && s->q_done.head != NULL) {
        sreq* q = q_pop(&s->q_done);
        q->state = RQ_POLLED;
        tcx_completion* c = &out[(*n)++];
        memset(c, 0, sizeof(*c));
        c->id = q->id;
        c->user = q->r.user;
        c->kind = q->r.kind;
        c->status = q->status;
        c->packet_size = q->packet_size;
        c->batch_frames = q->batch_frames;
        if (q->r.kind == TCX_REQ_ENCODE) {
            c->dev_packet = q->out_pkt;
        } else {
            for (uint32_t p = 0; p < q->nplanes; ++p) {
                c->dev_planes[p] = q->planes[p];
                c->strides[p] = q->pstride[p];
            }
        }
        c->t_submit_ns = q->t_submit;
        c->t_admit_ns = q->t_admit;
        c->t_launch_ns = q->t_launch;
        c->t_done_ns = q->t_done;
    }
    pthread_mutex_unlock(&s->mu);
    return TC_OK;
}

int32_t tcx_server_release(tcx_server* s, uint64_t id)
{
    if (s == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "server NULL"); }
    pthread_mutex_lock(&s->mu);
    sreq** pp = &s->live;
    while (*pp != NULL && (*pp)->id != id) { pp = &(*pp)->live_next; }
    sreq* q = *pp;
    if (q == NULL || q->state != RQ_POLLED) {
        pthread_mutex_unlock(&s->mu);
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "request %llu is not a polled completion", (unsigned long long)id);
    }
    *pp = q->live_next;
    if (q->len != 0u) { (void)tc_arena_free(&s->arena, q->off, q->len); }
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    req_free(q);
    return TC_OK;
}

int32_t tcx_server_get_stats(tcx_server* s, tcx_server_stats* st)
{
    if (s == NULL || st == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "server or stats NULL"); }
    pthread_mutex_lock(&s->mu);
    *st = s->st;
    st->held_bytes = s->arena.used;
    st->peak_bytes = s->arena.peak;
    st->queued = s->q_wait.n + s->q_enc.n + s->q_dec.n;
    st->inflight = s->nbusy;
    pthread_mutex_unlock(&s->mu);
    return TC_OK;
}

void tcx_server_destroy(tcx_server* s)
{
    if (s == NULL) { return; }
    pthread_mutex_lock(&s->mu);
    s->stop = 1;
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->mu);
    if (s->started) { pthread_join(s->thr, NULL); }
    while (s->live != NULL) {
        sreq* q = s->live;
        s->live = q->live_next;
        req_free(q);
    }
    (void)tcg_free(s->base);
    tc_arena_fini(&s->arena);
    pthread_mutex_destroy(&s->mu);
    pthread_cond_destroy(&s->cv_work);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}
