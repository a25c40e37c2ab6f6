/* Device engine: contexts of grow-only HBM buffers, pinned staging and one
 * default-priority stream each, plus the encode / decode launch sequences
 * behind the public calls. A call takes any free context, so calls from
 * different host threads run concurrently on the device; a context is used by
 * one call at a time. Every operation is split into submit (plan, uploads,
 * launches, result download queued) and finish (wait, post-process), so one
 * thread can keep an encode and a decode batch in flight on two contexts.
 *
 * Batch model: n frames of one call share one flat block index space, one
 * descriptor upload and one launch of every kernel, so a batch costs the same
 * number of launches as a single frame.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "../kernels/tc_math.h"

typedef struct dbuf { uint64_t p; size_t cap; } dbuf;

typedef struct enc_frame_plan {
    tc_fhdr fh;
    uint32_t slice_rows;
    uint32_t nslices;
    int32_t qp_eff[4];
} enc_frame_plan;

struct tc_ectx {
    pthread_mutex_t mu;
    tcg_ctx* g;
    dbuf in, q, dcm, scratch, pkt, desc, res, out, src;
    uint8_t* hdesc;    /* pinned staging for descriptor uploads */
    size_t hdesc_cap;
    uint8_t* hres;     /* pinned staging for result downloads */
    size_t hres_cap;
    /* operation in flight between submit and finish */
    enc_frame_plan* plan;
    size_t o_res, o_fo, o_fr;
    uint64_t* out_dev;
    uint32_t nslices;
    /* no-allocation context: buffers are fixed views of the static device
     * scratch and static host arrays; a need beyond them fails instead of
     * allocating */
    int noalloc;
    void* na_keys;          /* work-list keys (TC_NA_MAX_SLICES) */
    uint64_t* na_out_dev;   /* (TC_NA_MAX_FRAMES x 4 + 1) output plane addresses */
};

/* at most the hardware queues HIP gives a process on this machine class
 * (GPU_MAX_HW_QUEUES = 4): more streams would share queues and serialise */
#define TC_MAX_CTX 4u

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static int g_state = 0;       /* 0 uninitialised, 1 ready, -1 failed */
static int32_t g_init_rc = 0;
static char g_init_msg[256];
static tc_ectx g_ctx[TC_MAX_CTX];
static uint32_t g_nctx = 0, g_rr = 0;
static pthread_mutex_t g_tmu = PTHREAD_MUTEX_INITIALIZER;
static tcg_timing g_last_enc, g_last_dec;

static size_t al256(size_t v) { return (v + 255u) & ~(size_t)255u; }

static int32_t dbuf_need(tc_ectx* c, dbuf* b, size_t bytes)
{
    if (bytes <= b->cap) { return TC_OK; }
    if (c->noalloc) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "no-allocation decode: %zu bytes needed, the fixed buffer holds %zu",
                       bytes, b->cap);
    }
    if (b->p) { (void)tcg_sync(c->g); (void)tcg_free(b->p); b->p = 0; b->cap = 0; }
    size_t cap = bytes + bytes / 4u + 4096u;
    if (tcg_malloc(&b->p, cap) != 0) {
        b->p = 0;
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "device allocation of %zu bytes failed: %s", cap, tcg_last_error());
    }
    b->cap = cap;
    return TC_OK;
}

static int32_t host_need(tc_ectx* c, uint8_t** p, size_t* cap, size_t bytes)
{
    if (bytes <= *cap) { return TC_OK; }
    if (c->noalloc) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "no-allocation decode: %zu host bytes needed, the fixed buffer holds %zu",
                       bytes, *cap);
    }
    if (*p) { (void)tcg_sync(c->g); (void)tcg_host_free(*p); *p = NULL; *cap = 0; }
    size_t n = bytes + bytes / 4u + 4096u;
    if (tcg_host_alloc((void**)p, n) != 0) {
        *p = NULL;
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "pinned host allocation of %zu bytes failed", n);
    }
    *cap = n;
    return TC_OK;
}

/* host arrays of a decode plan: heap blocks, or the no-allocation context's
 * fixed arrays, which are never freed */
static void plan_free(const tc_ectx* c, void* p)
{
    if (!c->noalloc) { free(p); }
}

/* work list of the persistent rANS2 launch: slice indices by descending
 * block count (longest first), ties in batch order */
typedef struct work_key { uint32_t nblk, idx; } work_key;

static int work_key_cmp(const void* a, const void* b)
{
    const work_key* x = (const work_key*)a;
    const work_key* y = (const work_key*)b;
    if (x->nblk != y->nblk) { return x->nblk > y->nblk ? -1 : 1; }
    return x->idx < y->idx ? -1 : (x->idx > y->idx ? 1 : 0);
}

static int32_t work_list(work_key* keys, uint32_t n, tcg_work* w)
{
    qsort(keys, n, sizeof(work_key), work_key_cmp);
    uint32_t* order = (uint32_t*)(w + 1);
    w->count = n;
    for (uint32_t i = 0; i < n; ++i) { order[i] = keys[i].idx; }
    return TC_OK;
}

/* launch shapes of the persistent rANS2 kernels and the transform kernels
 * for a batch (tiles: the batch's 32-block transform tiles) */
static void entropy_shape(tcg_batch_info* info, int encode, uint32_t nframes, uint32_t tiles)
{
    tc_tune_entropy(encode, info->max_cols, info->ent_count, nframes, &info->ent_wpg, &info->ent_groups);
    info->tile_tpw = tc_tune_tiles(encode, info->max_cols, tiles);
}

/* colour slice entropy coding of a frame generation */
static uint32_t slice_mode(const tc_fhdr* fh)
{
    if (fh->major == 1u) { return TCG_MODE_RICE; }
    if (fh->major == 2u) { return TCG_MODE_VLC; }
    return TCG_MODE_RANS2;
}

/* encode scratch bound per 8x8 block of a slice */
static size_t slice_bytes_per_block(const tc_fhdr* fh, uint32_t p)
{
    if (tc_is_alpha(fh, p)) { return 2u * TCG_SLICE_BYTES_PER_BLOCK; }
    return slice_mode(fh) == TCG_MODE_RANS2 ? TCG_SLICE_BYTES_PER_BLOCK : TCG_BITS_BYTES_PER_BLOCK;
}

static int32_t engine_init_locked(void)
{
    if (g_state == 0) {
        int32_t rc = tcg_init();
        if (rc == 0) {
            static double cost[4097];
            uint32_t crc[256], x2n[32];
            cost[0] = 0.0;
            for (uint32_t f = 1; f <= 4096u; ++f) { cost[f] = 12.0 - tc_ctx_log2((double)f); }
            tc_crc_tables(crc, x2n);
            rc = tcg_set_tables(cost, crc, x2n);
            static tcg_vlc_books books;
            if (rc == 0 && tc_vlc_books_build(&books) != TC_OK) { rc = -2; }
            if (rc == 0) { rc = tcg_set_vlc(&books); }
        }
        if (rc != 0) {
            g_state = -1;
            g_init_rc = rc == -16 ? TC_ERR_NOT_IMPLEMENTED : TC_ERR_OUT_OF_MEMORY;
            strncpy(g_init_msg, tcg_last_error(), sizeof(g_init_msg) - 1u);
        } else {
            g_state = 1;
        }
    }
    return g_state == 1 ? TC_OK : tc_fail(g_init_rc, "GPU backend unavailable: %s", g_init_msg);
}

int32_t tc_engine_try_acquire(tc_ectx** out)
{
    pthread_mutex_lock(&g_mu);
    int32_t rc = engine_init_locked();
    if (rc != TC_OK) { pthread_mutex_unlock(&g_mu); return rc; }
    for (uint32_t i = 0; i < g_nctx; ++i) {
        if (pthread_mutex_trylock(&g_ctx[i].mu) == 0) {
            pthread_mutex_unlock(&g_mu);
            *out = &g_ctx[i];
            return TC_OK;
        }
    }
    if (g_nctx < TC_MAX_CTX) {
        tc_ectx* c = &g_ctx[g_nctx];
        memset(c, 0, sizeof(*c));
        if (tcg_ctx_create(&c->g) != 0) {
            pthread_mutex_unlock(&g_mu);
            return tc_fail(TC_ERR_OUT_OF_MEMORY, "stream creation failed: %s", tcg_last_error());
        }
        pthread_mutex_init(&c->mu, NULL);
        pthread_mutex_lock(&c->mu);
        g_nctx++;
        pthread_mutex_unlock(&g_mu);
        *out = c;
        return TC_OK;
    }
    pthread_mutex_unlock(&g_mu);
    *out = NULL;
    return 1;
}

int32_t tc_engine_acquire(tc_ectx** out)
{
    const int32_t rc = tc_engine_try_acquire(out);
    if (rc != 1) { return rc; }
    /* every context is in use: wait for one, round robin */
    pthread_mutex_lock(&g_mu);
    tc_ectx* c = &g_ctx[g_rr++ % g_nctx];
    pthread_mutex_unlock(&g_mu);
    pthread_mutex_lock(&c->mu);
    *out = c;
    return TC_OK;
}

void tc_engine_release(tc_ectx* c) { pthread_mutex_unlock(&c->mu); }

/* The no-allocation context: one per process, created on its first use (its
 * stream and events), launching eagerly (no graph instantiation), with the
 * device buffers q / desc / res fixed inside the static scratch
 * (tcg_noalloc_scratch) and the host staging in static arrays. Calls through
 * it are serialised by its lock; every other buffer has capacity 0, so a
 * decode that would need one (host packets or planes, target-size or
 * alpha-dropping decodes) fails with TC_ERR_LIMIT_EXCEEDED. */
static tc_ectx g_na_ctx;
static int g_na_ready = 0;
static uint8_t g_na_hdesc[TCG_NOALLOC_DESC_BYTES];
static uint8_t g_na_hres[TCG_NOALLOC_RES_BYTES];
static work_key g_na_keys[TC_NA_MAX_SLICES];
static uint64_t g_na_out_dev[TC_NA_MAX_FRAMES * 4u + 1u];

int32_t tc_engine_acquire_noalloc(tc_ectx** out)
{
    pthread_mutex_lock(&g_mu);
    int32_t rc = engine_init_locked();
    if (rc == TC_OK && !g_na_ready) {
        tc_ectx* c = &g_na_ctx;
        memset(c, 0, sizeof(*c));
        uint64_t base = 0u;
        if (tcg_noalloc_scratch(&base) != 0) {
            rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "no-allocation scratch: %s", tcg_last_error());
        } else if (tcg_ctx_create(&c->g) != 0) {
            rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "stream creation failed: %s", tcg_last_error());
        } else {
            tcg_ctx_set_eager(c->g, 1);
            c->noalloc = 1;
            c->q.p = base;
            c->q.cap = TCG_NOALLOC_Q_BYTES;
            c->desc.p = base + TCG_NOALLOC_Q_BYTES;
            c->desc.cap = TCG_NOALLOC_DESC_BYTES;
            c->res.p = base + TCG_NOALLOC_Q_BYTES + TCG_NOALLOC_DESC_BYTES;
            c->res.cap = TCG_NOALLOC_RES_BYTES;
            c->hdesc = g_na_hdesc;
            c->hdesc_cap = sizeof(g_na_hdesc);
            c->hres = g_na_hres;
            c->hres_cap = sizeof(g_na_hres);
            c->na_keys = g_na_keys;
            c->na_out_dev = g_na_out_dev;
            pthread_mutex_init(&c->mu, NULL);
            g_na_ready = 1;
        }
    }
    pthread_mutex_unlock(&g_mu);
    if (rc != TC_OK) { return rc; }
    pthread_mutex_lock(&g_na_ctx.mu);
    *out = &g_na_ctx;
    return TC_OK;
}

int32_t tc_engine_staging(tc_ectx* c, size_t in_bytes, size_t out_bytes, uint64_t* in, uint64_t* out)
{
    int32_t rc;
    if ((rc = dbuf_need(c, &c->in, in_bytes ? in_bytes : 256u)) != TC_OK ||
        (rc = dbuf_need(c, &c->out, out_bytes ? out_bytes : 256u)) != TC_OK) {
        return rc;
    }
    *in = c->in.p;
    *out = c->out.p;
    return TC_OK;
}

size_t tc_engine_encode_cap(const topos_frame_config* cfg)
{
    tc_fhdr fh;
    tc_cfg_to_fhdr(cfg, &fh);
    if (tc_fhdr_geometry(&fh) != TC_OK) { return 0u; }
    const uint32_t sr = tc_slice_rows_of(cfg);
    uint32_t bands[4], total = 0;
    if (sr == 0u || tc_band_plan(&fh, sr, bands, &total) != TC_OK) { return 0u; }
    size_t cap = TC_HDR_SIZE;
    for (uint32_t p = 0; p < fh.plane_count; ++p) {
        for (uint32_t y0 = 0; y0 < fh.brows[p]; y0 += sr) {
            const uint32_t h = fh.brows[p] - y0 < sr ? fh.brows[p] - y0 : sr;
            cap += TC_SLICE_HDR_SIZE + (size_t)h * fh.bcols[p] * slice_bytes_per_block(&fh, p) + TCG_SLICE_BYTES_EXTRA;
        }
    }
    return cap < TC_MAX_PACKET ? cap : TC_MAX_PACKET;
}

tcg_ctx* tc_engine_stream(tc_ectx* c) { return c->g; }

void tc_engine_graph_stats(uint64_t* hits, uint64_t* builds)
{
    uint64_t h = 0, b = 0;
    pthread_mutex_lock(&g_mu);
    const uint32_t n = g_nctx;
    pthread_mutex_unlock(&g_mu);
    for (uint32_t i = 0; i < n; ++i) {
        uint64_t a = 0, c = 0;
        tcg_graph_stats(g_ctx[i].g, &a, &c);
        h += a;
        b += c;
    }
    *hits = h;
    *builds = b;
}

void tc_engine_last_timing(tcg_timing* enc, tcg_timing* dec)
{
    pthread_mutex_lock(&g_tmu);
    *enc = g_last_enc;
    *dec = g_last_dec;
    pthread_mutex_unlock(&g_tmu);
}

static float g_last_cvt_ms;

void tc_engine_record_convert(tc_ectx* c)
{
    if (!tcg_timing_on()) { return; }
    tcg_timing t;
    tcg_ctx_timing(c->g, &t);
    pthread_mutex_lock(&g_tmu);
    g_last_cvt_ms = t.convert_ms;
    pthread_mutex_unlock(&g_tmu);
}

float tc_engine_last_convert_ms(void)
{
    pthread_mutex_lock(&g_tmu);
    const float v = g_last_cvt_ms;
    pthread_mutex_unlock(&g_tmu);
    return v;
}

static void record_timing(tc_ectx* c, int encode)
{
    if (!tcg_timing_on()) { return; }
    tcg_timing t;
    tcg_ctx_timing(c->g, &t);
    pthread_mutex_lock(&g_tmu);
    if (encode) { g_last_enc = t; } else { g_last_dec = t; }
    pthread_mutex_unlock(&g_tmu);
}

/* ------------------------------------------------------------------ encode */

int32_t tc_engine_encode_submit(tc_ectx* c, tc_enc_req* reqs, uint32_t n)
{
    c->plan = NULL;
    enc_frame_plan* plan = (enc_frame_plan*)calloc(n ? n : 1u, sizeof(enc_frame_plan));
    if (plan == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "encode plan"); }
    uint32_t nplanes = 0, nslices = 0, nblocks = 0, ntiles = 0, nqctx = 0;
    size_t in_bytes = 0, scratch_bytes = 0, pkt_bytes = 0;
    int32_t rc = TC_OK;
    for (uint32_t f = 0; f < n && rc == TC_OK; ++f) {
        tc_enc_req* r = &reqs[f];
        enc_frame_plan* P = &plan[f];
        tc_cfg_to_fhdr(r->cfg, &P->fh);
        (void)tc_fhdr_geometry(&P->fh);
        P->slice_rows = tc_slice_rows_of(r->cfg);
        uint32_t bands[4];
        rc = tc_band_plan(&P->fh, P->slice_rows, bands, &P->nslices);
        if (rc != TC_OK) { break; }
        P->fh.slice_count = (uint16_t)P->nslices;
        for (uint32_t p = 0; p < P->fh.plane_count; ++p) {
            const tc_fhdr* fh = &P->fh;
            nplanes++;
            nqctx++;
            if (!tc_is_alpha(fh, p)) {
                P->qp_eff[p] = tc_qp_eff(fh, r->cfg, p);
                const uint32_t pt = (uint32_t)fh->brows[p] * ((fh->bcols[p] + TCG_FQ_BLOCKS - 1u) / TCG_FQ_BLOCKS);
                ntiles = pt > ntiles ? pt : ntiles;
            }
            nblocks += (uint32_t)fh->bcols[p] * fh->brows[p];
            if (r->dev_planes == NULL) {
                const size_t stride = ((size_t)fh->pvw[p] + 7u) & ~(size_t)7u;
                in_bytes += al256(stride * fh->pvh[p] * 2u);
            }
            for (uint32_t y0 = 0; y0 < fh->brows[p]; y0 += P->slice_rows) {
                const uint32_t h = fh->brows[p] - y0 < P->slice_rows ? fh->brows[p] - y0 : P->slice_rows;
                const size_t cap = (size_t)h * fh->bcols[p] * slice_bytes_per_block(fh, p) + TCG_SLICE_BYTES_EXTRA;
                scratch_bytes += al256(cap);
                pkt_bytes += TC_SLICE_HDR_SIZE + cap;
            }
        }
        nslices += P->nslices;
        pkt_bytes = al256(pkt_bytes + TC_HDR_SIZE + 1024u);
    }
    if (rc != TC_OK) { free(plan); return rc; }

    /* descriptor blob: qctx | planes | slices | frames ; results | frame_out */
    const size_t o_q = 0;
    const size_t o_pl = al256(o_q + (size_t)nqctx * sizeof(tcg_qctx));
    const size_t o_sl = al256(o_pl + (size_t)nplanes * sizeof(tcg_eplane));
    const size_t o_fr = al256(o_sl + (size_t)nslices * sizeof(tcg_eslice));
    const size_t o_wk = al256(o_fr + (size_t)n * sizeof(tcg_eframe));
    const size_t desc_bytes = al256(o_wk + sizeof(tcg_work) + (size_t)nslices * sizeof(uint32_t));
    const size_t o_res = 0;
    const size_t o_fo = al256((size_t)nslices * sizeof(tcg_eresult));
    const size_t res_bytes = al256(o_fo + (size_t)n * sizeof(tcg_eframe_out));

    if ((rc = dbuf_need(c, &c->q, (size_t)nblocks * 64u * 4u)) != TC_OK ||
        (rc = dbuf_need(c, &c->dcm, (size_t)nblocks * 4u + 256u)) != TC_OK ||
        (rc = dbuf_need(c, &c->scratch, scratch_bytes)) != TC_OK ||
        (rc = dbuf_need(c, &c->desc, desc_bytes)) != TC_OK ||
        (rc = dbuf_need(c, &c->res, res_bytes)) != TC_OK ||
        (rc = host_need(c, &c->hdesc, &c->hdesc_cap, desc_bytes)) != TC_OK ||
        (rc = host_need(c, &c->hres, &c->hres_cap, res_bytes)) != TC_OK) {
        free(plan);
        return rc;
    }
    int need_pkt = 0;
    for (uint32_t f = 0; f < n; ++f) { if (reqs[f].dev_out == 0u) { need_pkt = 1; } }
    if (need_pkt && (rc = dbuf_need(c, &c->pkt, pkt_bytes)) != TC_OK) { free(plan); return rc; }
    if (in_bytes && (rc = dbuf_need(c, &c->in, in_bytes)) != TC_OK) { free(plan); return rc; }

    memset(c->hdesc, 0, desc_bytes);
    tcg_qctx* hq = (tcg_qctx*)(c->hdesc + o_q);
    tcg_eplane* hp = (tcg_eplane*)(c->hdesc + o_pl);
    tcg_eslice* hs = (tcg_eslice*)(c->hdesc + o_sl);
    tcg_eframe* hf = (tcg_eframe*)(c->hdesc + o_fr);
    uint32_t ip = 0, is = 0, iq = 0, blk = 0, tile = 0;
    size_t in_off = 0, scr_off = 0, pkt_off = 0;
    tcg_batch_info info;
    memset(&info, 0, sizeof(info));
    info.max_cols = 1u;
    work_key* wkeys = (work_key*)malloc(sizeof(work_key) * (nslices ? nslices : 1u));
    if (wkeys == NULL) { free(plan); return tc_fail(TC_ERR_OUT_OF_MEMORY, "work list"); }
    for (uint32_t f = 0; f < n && rc == TC_OK; ++f) {
        tc_enc_req* r = &reqs[f];
        enc_frame_plan* P = &plan[f];
        const tc_fhdr* fh = &P->fh;
        const uint32_t tbl = (fh->flags & 8u) ? 1u : 0u;
        const uint32_t s0 = is;
        size_t frame_cap = TC_HDR_SIZE;
        for (uint32_t p = 0; p < fh->plane_count; ++p) {
            tcg_eplane* E = &hp[ip];
            const uint32_t vw = fh->pvw[p], vh = fh->pvh[p];
            if (r->dev_planes != NULL) {
                E->src = r->dev_planes[p];
                E->stride = (uint32_t)(r->dev_strides && r->dev_strides[p] ? r->dev_strides[p] : vw);
            } else {
                const size_t stride = ((size_t)vw + 7u) & ~(size_t)7u;
                const size_t sst = r->input->strides[p] ? r->input->strides[p] : vw;
                E->src = c->in.p + in_off;
                E->stride = (uint32_t)stride;
                rc = tcg_h2d_2d(c->g, E->src, stride * 2u, r->input->planes[p], sst * 2u, (size_t)vw * 2u, vh);
                if (rc != 0) { rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "input upload: %s", tcg_last_error()); break; }
                in_off += al256(stride * vh * 2u);
            }
            E->vis_w = vw; E->vis_h = vh;
            E->bcols = fh->bcols[p]; E->brows = fh->brows[p];
            E->blk0 = blk;
            E->coded_w = fh->pcw[p];
            E->tile0 = tile;
            tc_qdiv32_consts((fh->bcols[p] + TCG_FQ_BLOCKS - 1u) / TCG_FQ_BLOCKS, &E->tile_m, &E->tile_l);
            /* quantiser context i belongs to plane i of the batch */
            E->qctx = iq;
            if (tc_is_alpha(fh, p)) {
                E->kind = 1u;
                E->alpha_shift = fh->alpha_mode == 2u ? 16u - fh->alpha_bit_depth : 0u;
                info.kinds |= TCG_HAS_ALPHA;
            } else {
                E->kind = 0u;
                info.kinds |= slice_mode(fh) == TCG_MODE_RANS2 ? TCG_HAS_RANS2 : TCG_HAS_BITS;
                info.max_cols = fh->bcols[p] > info.max_cols ? fh->bcols[p] : info.max_cols;
                tc_qctx_build(&hq[iq], tc_qm_tab[fh->qmatrix_id][tc_is_chroma(fh, p) ? 1 : 0],
                              (uint32_t)P->qp_eff[p], fh->bit_depth, tbl);
                tile += (uint32_t)fh->brows[p] * ((fh->bcols[p] + TCG_FQ_BLOCKS - 1u) / TCG_FQ_BLOCKS);
            }
            iq++;
            for (uint32_t y0 = 0; y0 < fh->brows[p]; y0 += P->slice_rows) {
                const uint32_t h = fh->brows[p] - y0 < P->slice_rows ? fh->brows[p] - y0 : P->slice_rows;
                tcg_eslice* S = &hs[is++];
                S->plane = ip; S->frame = f; S->y0 = y0; S->h = h;
                S->blk0 = blk + y0 * fh->bcols[p];
                S->nblk = h * fh->bcols[p];
                S->scratch = scr_off;
                S->scratch_cap = (uint32_t)(S->nblk * slice_bytes_per_block(fh, p) + TCG_SLICE_BYTES_EXTRA);
                S->mode = slice_mode(fh);
                S->qpd = tc_is_alpha(fh, p) ? 64u : (uint32_t)(64 + P->qp_eff[p] - (int32_t)fh->qp_base);
                S->hdr_plane = p;
                if (!tc_is_alpha(fh, p) && S->mode == TCG_MODE_RANS2) {
                    wkeys[info.ent_count].nblk = S->nblk;
                    wkeys[info.ent_count].idx = is - 1u;
                    info.ent_count++;
                }
                scr_off += al256(S->scratch_cap);
                frame_cap += TC_SLICE_HDR_SIZE + S->scratch_cap;
            }
            blk += (uint32_t)fh->bcols[p] * fh->brows[p];
            ip++;
        }
        if (rc != TC_OK) { break; }
        tcg_eframe* F = &hf[f];
        if (r->dev_out != 0u) {
            F->packet = r->dev_out;
            F->cap = (uint32_t)(r->dev_out_cap < TC_MAX_PACKET ? r->dev_out_cap : TC_MAX_PACKET);
        } else {
            F->packet = c->pkt.p + pkt_off;
            F->cap = (uint32_t)(frame_cap < TC_MAX_PACKET ? frame_cap : TC_MAX_PACKET);
            pkt_off = al256(pkt_off + frame_cap + 1024u);
        }
        F->s0 = s0; F->s1 = is;
        tc_fhdr tmp = *fh;
        tmp.packet_size = 0u;
        tc_fhdr_pack(&tmp, F->hdr);
    }
    if (rc != TC_OK) { free(wkeys); free(plan); return rc; }
    (void)work_list(wkeys, info.ent_count, (tcg_work*)(c->hdesc + o_wk));
    free(wkeys);
    info.work = (tcg_work*)(uintptr_t)(c->desc.p + o_wk);
    entropy_shape(&info, 1, n, tile);
    if (tcg_h2d(c->g, c->desc.p, c->hdesc, desc_bytes) != 0) {
        free(plan);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "descriptor upload: %s", tcg_last_error());
    }
    if (tcg_encode(c->g, (const tcg_qctx*)(uintptr_t)(c->desc.p + o_q), nplanes,
                   (const tcg_eplane*)(uintptr_t)(c->desc.p + o_pl), nslices,
                   (const tcg_eslice*)(uintptr_t)(c->desc.p + o_sl), n,
                   (const tcg_eframe*)(uintptr_t)(c->desc.p + o_fr), c->q.p, c->dcm.p, ntiles, c->scratch.p,
                   (tcg_eresult*)(uintptr_t)(c->res.p + o_res),
                   (tcg_eframe_out*)(uintptr_t)(c->res.p + o_fo), &info) != 0) {
        free(plan);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "encode launch failed: %s", tcg_last_error());
    }
    if (tcg_d2h(c->g, c->hres, c->res.p, res_bytes) != 0) {
        free(plan);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "encode result download: %s", tcg_last_error());
    }
    c->plan = plan;
    c->o_res = o_res;
    c->o_fo = o_fo;
    c->o_fr = o_fr;
    return TC_OK;
}

int32_t tc_engine_encode_finish(tc_ectx* c, tc_enc_req* reqs, uint32_t n)
{
    enc_frame_plan* plan = c->plan;
    c->plan = NULL;
    if (plan == NULL) { return tc_fail(TC_ERR_STATE, "no encode in flight"); }
    if (tcg_sync(c->g) != 0) {
        free(plan);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "encode: %s", tcg_last_error());
    }
    record_timing(c, 1);
    const tcg_eframe* hf = (const tcg_eframe*)(c->hdesc + c->o_fr);
    const tcg_eframe_out* fo = (const tcg_eframe_out*)(c->hres + c->o_fo);
    int32_t first = TC_OK;
    for (uint32_t f = 0; f < n; ++f) {
        tc_enc_req* r = &reqs[f];
        const tcg_eframe_out* o = &fo[f];
        r->packet_size = o->size;
        if (o->status != 0) {
            r->rc = o->status == -1 ? tc_fail(TC_ERR_INVALID_ARGUMENT,
                                              "token domain violated (bit_depth %u qp too low for full-scale content)",
                                              (unsigned)plan[f].fh.bit_depth)
                                    : tc_fail(o->status, "encode failed on device (status %d)", (int)o->status);
        } else {
            if (r->stats) {
                topos_frame_stats* st = r->stats;
                st->packet_size = o->size;
                st->color_payload_bytes = o->color_payload;
                st->alpha_payload_bytes = o->alpha_payload;
                st->color_header_bytes = TC_HDR_SIZE + TC_SLICE_HDR_SIZE * o->color_slices;
                st->alpha_header_bytes = TC_SLICE_HDR_SIZE * o->alpha_slices;
                st->slice_count = (uint16_t)plan[f].nslices;
                st->qp_base = plan[f].fh.qp_base;
                uint32_t merr = 0u;
                const tcg_eresult* rs = (const tcg_eresult*)(c->hres + c->o_res);
                for (uint32_t i = hf[f].s0; i < hf[f].s1; ++i) { merr = rs[i].aux > merr ? rs[i].aux : merr; }
                st->alpha_max_abs_error = (uint16_t)merr;
            }
            r->rc = TC_OK;
            if (r->out != NULL || r->dev_out == 0u) {
                if (o->size > r->out_cap) {
                    r->rc = tc_fail(TC_ERR_BUFFER_TOO_SMALL, "out_cap %zu < %u bytes needed", r->out_cap,
                                    (unsigned)o->size);
                } else if (r->out != NULL) {
                    if (tcg_d2h_sync(c->g, r->out, hf[f].packet, o->size) != 0) {
                        r->rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "packet download: %s", tcg_last_error());
                    }
                }
            }
        }
        if (first == TC_OK && r->rc != TC_OK) { first = r->rc; }
    }
    free(plan);
    return first;
}

int32_t tc_engine_encode(tc_ectx* c, tc_enc_req* reqs, uint32_t n)
{
    int32_t rc = tc_engine_encode_submit(c, reqs, n);
    return rc != TC_OK ? rc : tc_engine_encode_finish(c, reqs, n);
}

/* ------------------------------------------------------------------ decode */

void tc_target_plane_dims(const tc_fhdr* fh, uint32_t tw, uint32_t th, uint32_t p, uint32_t* w, uint32_t* h)
{
    if (fh->pixel_format == 3u) {
        *w = (tw + 1u) / 2u;
        *h = (th + 1u) / 2u;
    } else {
        *w = fh->pixel_format == 0u && (p == 1u || p == 2u) ? (tw + 1u) / 2u : tw;
        *h = th;
    }
}

/* output stride of plane p of a request (elements) */
static size_t dec_out_stride(const tc_dec_req* r, uint32_t p, uint32_t w)
{
    if (r->dev_planes != NULL) { return r->dev_strides && r->dev_strides[p] ? r->dev_strides[p] : w; }
    return w;
}

/* A target-size request decodes every plane at source size into `src`
 * (coefficients up to the request's limit, no deblocking), then a target
 * descriptor samples it into the output (through `tgt` and the deblock pass
 * on a 4-sample grid when the target is exactly 2:1 and the stream deblocks).
 * A dropped alpha plane decodes into `src` and is not output. */
int32_t tc_engine_decode_submit(tc_ectx* c, tc_dec_req* reqs, uint32_t n)
{
    c->out_dev = NULL;
    uint32_t nplanes = 0, nslices = 0, nblocks = 0, ntiles = 0, maxrows = 1, nscaled = 0;
    size_t pkt_bytes = 0, out_bytes = 0, src_bytes = 0;
    for (uint32_t f = 0; f < n; ++f) {
        const tc_dec_req* r = &reqs[f];
        const tc_fhdr* fh = &r->view->fh;
        const int scaled = r->tw != 0u;
        nscaled += scaled ? 1u : 0u;
        nslices += r->view->nslices;
        if (r->dev_data == 0u) { pkt_bytes += al256(r->size + 1024u); }
        const int db = (fh->flags & 4u) != 0u;
        for (uint32_t p = 0; p < fh->plane_count; ++p) {
            nplanes++;
            nblocks += (uint32_t)fh->bcols[p] * fh->brows[p];
            maxrows = fh->brows[p] > maxrows ? fh->brows[p] : maxrows;
            if (!tc_is_alpha(fh, p)) {
                const uint32_t pt = (uint32_t)fh->brows[p] * ((fh->bcols[p] + TCG_IQ_BLOCKS - 1u) / TCG_IQ_BLOCKS);
                ntiles = pt > ntiles ? pt : ntiles;
            }
            const int dropped = r->skip_alpha && tc_is_alpha(fh, p);
            uint32_t ow = fh->pvw[p], oh = fh->pvh[p];
            if (scaled) { tc_target_plane_dims(fh, r->tw, r->th, p, &ow, &oh); }
            if (scaled || dropped) { src_bytes += al256((size_t)fh->pvw[p] * fh->pvh[p] * 2u); }
            if (dropped) { continue; }
            const size_t ost = dec_out_stride(r, p, ow);
            if (r->dev_planes == NULL) { out_bytes += al256((size_t)ow * oh * 2u); }
        }
    }
    const size_t o_q = 0;
    const size_t o_pl = al256(o_q + (size_t)nslices * sizeof(tcg_qctx));
    const size_t o_sl = al256(o_pl + (size_t)nplanes * sizeof(tcg_dplane));
    const size_t o_rs = al256(o_sl + (size_t)nslices * sizeof(tcg_dslice));
    const size_t o_wk = al256(o_rs + (size_t)nplanes * maxrows * sizeof(tcg_rowrec));
    const size_t o_tg = al256(o_wk + sizeof(tcg_work) + (size_t)nslices * sizeof(uint32_t));
    const size_t desc_bytes = al256(o_tg + (nscaled ? (size_t)nplanes * sizeof(tcg_dplane) : 0u));
    const size_t res_bytes = al256((size_t)nslices * sizeof(int32_t));
    int32_t rc;
    if ((rc = dbuf_need(c, &c->q, (size_t)nblocks * 64u * 4u)) != TC_OK ||
        (rc = dbuf_need(c, &c->desc, desc_bytes)) != TC_OK ||
        (rc = dbuf_need(c, &c->res, res_bytes)) != TC_OK ||
        (rc = host_need(c, &c->hdesc, &c->hdesc_cap, desc_bytes)) != TC_OK ||
        (rc = host_need(c, &c->hres, &c->hres_cap, res_bytes)) != TC_OK) {
        return rc;
    }
    if (pkt_bytes && (rc = dbuf_need(c, &c->pkt, pkt_bytes)) != TC_OK) { return rc; }
    if (out_bytes && (rc = dbuf_need(c, &c->out, out_bytes)) != TC_OK) { return rc; }
    if (src_bytes && (rc = dbuf_need(c, &c->src, src_bytes)) != TC_OK) { return rc; }

    memset(c->hdesc, 0, desc_bytes);
    tcg_qctx* hq = (tcg_qctx*)(c->hdesc + o_q);
    tcg_dplane* hp = (tcg_dplane*)(c->hdesc + o_pl);
    tcg_dslice* hs = (tcg_dslice*)(c->hdesc + o_sl);
    tcg_rowrec* hrs = (tcg_rowrec*)(c->hdesc + o_rs);
    tcg_dplane* ht = nscaled ? (tcg_dplane*)(c->hdesc + o_tg) : NULL;
    uint32_t ip = 0, is = 0, iq = 0, blk = 0, tile = 0, row = 0, max_px = 0, tdbt = 0;
    size_t pkt_off = 0, out_off = 0, src_off = 0;
    tcg_batch_info info;
    memset(&info, 0, sizeof(info));
    info.max_cols = 1u;
    work_key* wkeys;
    uint64_t* out_dev;
    if (c->noalloc) {
        if (nslices > TC_NA_MAX_SLICES || n > TC_NA_MAX_FRAMES) {
            return tc_fail(TC_ERR_LIMIT_EXCEEDED, "no-allocation decode: %u frames / %u slices exceed %u / %u", n,
                           nslices, TC_NA_MAX_FRAMES, TC_NA_MAX_SLICES);
        }
        wkeys = (work_key*)c->na_keys;
        out_dev = c->na_out_dev;
        memset(out_dev, 0, ((size_t)n * 4u + 1u) * sizeof(uint64_t));
    } else {
        wkeys = (work_key*)malloc(sizeof(work_key) * (nslices ? nslices : 1u));
        if (wkeys == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "work list"); }
        out_dev = (uint64_t*)calloc((size_t)n * 4u + 1u, sizeof(uint64_t));
        if (out_dev == NULL) { free(wkeys); return tc_fail(TC_ERR_OUT_OF_MEMORY, "decode plan"); }
    }
    for (uint32_t f = 0; f < n; ++f) {
        tc_dec_req* r = &reqs[f];
        const tc_view* v = r->view;
        const tc_fhdr* fh = &v->fh;
        const int scaled = r->tw != 0u;
        uint64_t pk = r->dev_data;
        if (pk == 0u) {
            pk = c->pkt.p + pkt_off;
            if (tcg_h2d(c->g, pk, r->data, r->size) != 0) {
                plan_free(c, wkeys);
                plan_free(c, out_dev);
                return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet upload: %s", tcg_last_error());
            }
            pkt_off += al256(r->size + 1024u);
        }
        const uint32_t tbl = (fh->flags & 8u) ? 1u : 0u;
        const int db = (fh->flags & 4u) != 0u;
        const int half = scaled && r->tw * 2u == fh->visible_w && r->th * 2u == fh->visible_h;
        const uint32_t ip0 = ip;
        for (uint32_t p = 0; p < fh->plane_count; ++p) {
            tcg_dplane* D = &hp[ip + p];
            const uint32_t vw = fh->pvw[p], vh = fh->pvh[p];
            const int dropped = r->skip_alpha && tc_is_alpha(fh, p);
            uint32_t ow = vw, oh = vh;
            if (scaled) { tc_target_plane_dims(fh, r->tw, r->th, p, &ow, &oh); }
            const uint32_t ost = (uint32_t)dec_out_stride(r, p, ow);
            uint64_t final_dst = 0u;
            if (!dropped) {
                if (r->dev_planes != NULL) {
                    final_dst = r->dev_planes[p];
                } else {
                    final_dst = c->out.p + out_off;
                    out_off += al256((size_t)ow * oh * 2u);
                }
                out_dev[(size_t)f * 4u + p] = final_dst;
            }
            if (scaled || dropped) {
                D->dst = c->src.p + src_off;
                D->stride = vw;
                src_off += al256((size_t)vw * vh * 2u);
            } else {
                D->dst = final_dst;
                D->stride = ost;
            }
            D->vis_w = vw; D->vis_h = vh;
            D->bcols = fh->bcols[p]; D->brows = fh->brows[p];
            D->blk0 = blk;
            D->row_slice = row;
            D->coded_w = fh->pcw[p];
            D->tile0 = tile;
            tc_qdiv32_consts((fh->bcols[p] + TCG_IQ_BLOCKS - 1u) / TCG_IQ_BLOCKS, &D->tile_m, &D->tile_l);
            D->limit = scaled ? r->limit : 63u;
            const int colour = !tc_is_alpha(fh, p);
            if (!colour) {
                D->kind = 1u;
                info.kinds |= TCG_HAS_ALPHA;
                D->fill = 0xFFFFu;
                D->maxv = 0xFFFFu;
                D->alpha_mask = fh->alpha_mode == 2u ? (1u << (16u - fh->alpha_bit_depth)) - 1u : 0u;
            } else {
                D->kind = 0u;
                info.kinds |= slice_mode(fh) == TCG_MODE_RANS2 ? TCG_HAS_RANS2 : TCG_HAS_BITS;
                info.max_cols = fh->bcols[p] > info.max_cols ? fh->bcols[p] : info.max_cols;
                D->fill = 1u << (fh->bit_depth - 1u);
                D->maxv = (1u << fh->bit_depth) - 1u;
                tile += (uint32_t)fh->brows[p] * ((fh->bcols[p] + TCG_IQ_BLOCKS - 1u) / TCG_IQ_BLOCKS);
                if (db && p < 3u && !scaled && vw % 8u == 0u && vh % 8u == 0u && vw >= 16u && vh >= 16u &&
                    fh->qp_base <= 95u && fh->bit_depth >= 8u && fh->bit_depth <= 16u) {
                    D->deblock = 1u;
                    D->db_unit = 8u;
                    D->db_alpha = tc_deblock_alpha(tbl, fh->qp_base);
                }
            }
            if (ht != NULL) {
                tcg_dplane* T = &ht[ip + p];
                if (scaled && !dropped) {
                    T->sample = 1u;
                    T->dst = final_dst;
                    T->stride = ost;
                    T->vis_w = ow; T->vis_h = oh;
                    T->maxv = D->maxv;
                    T->kind = D->kind;
                    max_px = ow * oh > max_px ? ow * oh : max_px;
                    if (colour && db && p < 3u && half && ow % 4u == 0u && oh % 4u == 0u && ow >= 8u &&
                        oh >= 8u && fh->qp_base <= 95u && fh->bit_depth >= 8u && fh->bit_depth <= 16u) {
                        T->deblock = 1u;
                        T->db_unit = 4u;
                        T->db_alpha = tc_deblock_alpha(tbl, fh->qp_base);
                        tdbt = tcg_db_grid_max(tdbt, tcg_db_grid(ow, oh, 4u));
                    }
                }
            }
            blk += (uint32_t)fh->bcols[p] * fh->brows[p];
            row += fh->brows[p];
        }
        /* slices; quantiser contexts deduplicated per (plane class, qp) within the frame */
        int32_t qkey[16];
        uint32_t qidx[16], nk = 0;
        for (uint32_t i = 0; i < v->nslices; ++i) {
            const tc_shdr* h = &v->sl[i];
            tcg_dslice* S = &hs[is];
            const uint32_t p = h->plane;
            const tcg_dplane* D = &hp[ip0 + p];
            S->payload = pk + v->off[i];
            S->size = h->size;
            S->crc = h->crc;
            S->plane = ip0 + p;
            S->y0 = h->y0; S->h = h->h;
            S->blk0 = D->blk0 + (uint32_t)h->y0 * fh->bcols[p];
            S->k1 = h->k1; S->k2 = h->k2; S->k3 = h->k3;
            S->mode = slice_mode(fh);
            if (!tc_is_alpha(fh, p) && S->mode == TCG_MODE_RANS2) {
                wkeys[info.ent_count].nblk = (uint32_t)h->h * fh->bcols[p];
                wkeys[info.ent_count].idx = is;
                info.ent_count++;
            }
            S->dc_max = fh->bit_depth > 12u ? 0x0FFFFFFFu : 0x08000000u;
            S->lv_max = fh->bit_depth > 12u ? 0x07FFFFFFu : 0x04000000u;
            if (tc_is_alpha(fh, p)) {
                S->kind = 1u;
                S->qctx = 0u;
            } else {
                S->kind = 0u;
                const int32_t qp = (int32_t)fh->qp_base + (int32_t)h->qpd - 64;
                const int32_t key = qp * 2 + (tc_is_chroma(fh, p) ? 1 : 0);
                uint32_t k = 0;
                while (k < nk && qkey[k] != key) { ++k; }
                if (k == nk) {
                    if (nk < 16u) { qkey[nk] = key; qidx[nk] = iq; nk++; }
                    tc_qctx_build(&hq[iq], tc_qm_tab[fh->qmatrix_id][tc_is_chroma(fh, p) ? 1 : 0], (uint32_t)qp,
                                  fh->bit_depth, tbl);
                    S->qctx = iq++;
                } else {
                    S->qctx = qidx[k];
                }
            }
            for (uint32_t y = 0; y < h->h; ++y) {
                tcg_rowrec* R = &hrs[(size_t)(ip0 + p) * maxrows + h->y0 + y];
                R->qctx = S->qctx;
                R->slice = is;
            }
            is++;
        }
        ip += fh->plane_count;
    }
    /* deblock grid extent over the filtered planes */
    uint32_t dbt = 0;
    for (uint32_t i = 0; i < ip; ++i) {
        if (hp[i].deblock) { dbt = tcg_db_grid_max(dbt, tcg_db_grid(hp[i].vis_w, hp[i].vis_h, 8u)); }
    }
    (void)work_list(wkeys, info.ent_count, (tcg_work*)(c->hdesc + o_wk));
    plan_free(c, wkeys);
    info.work = (tcg_work*)(uintptr_t)(c->desc.p + o_wk);
    entropy_shape(&info, 0, n, tile);
    if (tcg_h2d(c->g, c->desc.p, c->hdesc, desc_bytes) != 0) {
        plan_free(c, out_dev);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "descriptor upload: %s", tcg_last_error());
    }
    if (tcg_decode(c->g, (const tcg_qctx*)(uintptr_t)(c->desc.p + o_q), nplanes,
                   (const tcg_dplane*)(uintptr_t)(c->desc.p + o_pl), nslices,
                   (const tcg_dslice*)(uintptr_t)(c->desc.p + o_sl),
                   (const tcg_rowrec*)(uintptr_t)(c->desc.p + o_rs), maxrows, c->q.p, ntiles,
                   (int32_t*)(uintptr_t)c->res.p, dbt, &info) != 0) {
        plan_free(c, out_dev);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "decode launch failed: %s", tcg_last_error());
    }
    if (ht != NULL && tcg_decode_sample(c->g, (const tcg_dplane*)(uintptr_t)(c->desc.p + o_pl),
                                        (const tcg_dplane*)(uintptr_t)(c->desc.p + o_tg), nplanes, max_px,
                                        tdbt) != 0) {
        plan_free(c, out_dev);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "target-size sample launch failed: %s", tcg_last_error());
    }
    if (tcg_d2h(c->g, c->hres, c->res.p, (size_t)nslices * sizeof(int32_t)) != 0) {
        plan_free(c, out_dev);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "status download: %s", tcg_last_error());
    }
    c->out_dev = out_dev;
    c->nslices = nslices;
    return TC_OK;
}

int32_t tc_engine_decode_finish(tc_ectx* c, tc_dec_req* reqs, uint32_t n)
{
    uint64_t* out_dev = c->out_dev;
    c->out_dev = NULL;
    if (out_dev == NULL) { return tc_fail(TC_ERR_STATE, "no decode in flight"); }
    if (tcg_sync(c->g) != 0) {
        plan_free(c, out_dev);
        return tc_fail(TC_ERR_OUT_OF_MEMORY, "decode: %s", tcg_last_error());
    }
    record_timing(c, 0);
    const int32_t* st = (const int32_t*)c->hres;
    int32_t first = TC_OK;
    uint32_t sbase = 0;
    for (uint32_t f = 0; f < n; ++f) {
        tc_dec_req* r = &reqs[f];
        const tc_view* v = r->view;
        const tc_fhdr* fh = &v->fh;
        uint32_t conc = 0;
        for (uint32_t i = 0; i < v->nslices; ++i) {
            /* a dropped alpha plane's slices are not decoded for the caller: status OK */
            const int dropped = r->skip_alpha && tc_is_alpha(fh, v->sl[i].plane);
            const uint8_t s = st[sbase + i] != 0 && !dropped ? TC_FRAME_SLICE_CONCEALED : TC_FRAME_SLICE_OK;
            if (r->info) { r->info->slice_status[i] = s; }
            conc += s;
        }
        sbase += v->nslices;
        if (r->info) { r->info->concealed_slices = (uint16_t)conc; }
        r->rc = conc ? tc_fail(TC_WARN_CONCEALED, "%u of %u slices concealed", conc, (unsigned)v->nslices) : TC_OK;
        if (r->dev_planes == NULL && r->planes != NULL) {
            for (uint32_t p = 0; p < fh->plane_count; ++p) {
                if (r->skip_alpha && tc_is_alpha(fh, p)) { continue; }
                uint32_t ow = fh->pvw[p], oh = fh->pvh[p];
                if (r->tw != 0u) { tc_target_plane_dims(fh, r->tw, r->th, p, &ow, &oh); }
                const size_t dst_stride = r->strides && r->strides[p] ? r->strides[p] : ow;
                if (tcg_d2h_2d(c->g, r->planes[p], dst_stride * 2u, out_dev[(size_t)f * 4u + p], (size_t)ow * 2u,
                               (size_t)ow * 2u, oh) != 0) {
                    r->rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "plane download: %s", tcg_last_error());
                    break;
                }
            }
        }
        if (first == TC_OK && r->rc != TC_OK) { first = r->rc; }
    }
    plan_free(c, out_dev);
    return first;
}

int32_t tc_engine_decode(tc_ectx* c, tc_dec_req* reqs, uint32_t n)
{
    int32_t rc = tc_engine_decode_submit(c, reqs, n);
    return rc != TC_OK ? rc : tc_engine_decode_finish(c, reqs, n);
}
