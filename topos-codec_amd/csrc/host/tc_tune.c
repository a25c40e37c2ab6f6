/* Autotuner for the persistent rANS2 entropy launches.
 *
 * Knobs per launch: waves per workgroup (1, 2 or 4: slices a workgroup works
 * on at once, each wave on its own slice) and the grid (one wave per slice,
 * or only as many workgroups as fit on the device at once, the waves then
 * taking further slices from the work list). Neither changes a byte of the
 * output. The best pair is kept per (direction, widest plane in blocks,
 * number of rANS2 slices rounded up to a power of two): the resolution and
 * profile of a batch decide both. tcx_autotune_batch times every candidate on
 * a caller's representative batch with HIP events around the entropy kernel;
 * TOPOS_TUNE_FILE names a text file the table is loaded from at start-up and
 * saved to after each tuning run.
 */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tc_host.h"
#include "topos_gpu.h"

#define TUNE_MAX 64u

typedef struct tune_ent {
    uint32_t encode, cols, bucket;
    uint32_t wpg, mode;  /* mode 0: one wave per slice, 1: device-sized persistent grid */
    float ms;
    uint64_t used;
} tune_ent;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static tune_ent g_tab[TUNE_MAX];
static uint32_t g_n = 0;
static uint64_t g_clock = 0;
static int g_loaded = 0;

/* a forced shape for the calling thread (tuning runs), and the key of the
 * calling thread's last launch */
static __thread int t_force = 0;
static __thread uint32_t t_wpg = 1u, t_mode = 0u, t_groups = 0u;
static __thread uint32_t t_last_cols = 0u, t_last_count = 0u;
static __thread uint32_t t_tpw = 0u;  /* forced transform tiles per workgroup (0: tuned) */
static __thread uint32_t t_last_tiles = 0u;

/* table kinds (the `encode` field): 0 decode entropy, 1 encode entropy,
 * 2 FDCT tiles per workgroup, 3 IDCT tiles per workgroup */
#define KIND_FDCT 2u
#define KIND_IDCT 3u

static uint32_t bucket_of(uint32_t n)
{
    uint32_t b = 1u;
    while (b < n && b < 0x80000000u) { b <<= 1; }
    return b;
}

static void load_locked(void)
{
    if (g_loaded) { return; }
    g_loaded = 1;
    const char* path = getenv("TOPOS_TUNE_FILE");
    if (path == NULL || path[0] == '\0') { return; }
    FILE* f = fopen(path, "r");
    if (f == NULL) { return; }
    tune_ent e;
    memset(&e, 0, sizeof(e));
    while (g_n < TUNE_MAX && fscanf(f, "%u %u %u %u %u %f", &e.encode, &e.cols, &e.bucket, &e.wpg, &e.mode, &e.ms) == 6) {
        if ((e.wpg == 1u || e.wpg == 2u || e.wpg == 4u) && e.mode <= 1u && e.encode <= KIND_IDCT) { g_tab[g_n++] = e; }
    }
    fclose(f);
}

static void save_locked(void)
{
    const char* path = getenv("TOPOS_TUNE_FILE");
    if (path == NULL || path[0] == '\0') { return; }
    FILE* f = fopen(path, "w");
    if (f == NULL) { return; }
    for (uint32_t i = 0; i < g_n; ++i) {
        fprintf(f, "%u %u %u %u %u %.4f\n", g_tab[i].encode, g_tab[i].cols, g_tab[i].bucket, g_tab[i].wpg,
                g_tab[i].mode, (double)g_tab[i].ms);
    }
    fclose(f);
}

static void record(uint32_t encode, uint32_t cols, uint32_t count, uint32_t wpg, uint32_t mode, float ms)
{
    pthread_mutex_lock(&g_mu);
    load_locked();
    const uint32_t b = bucket_of(count);
    uint32_t slot = g_n;
    for (uint32_t i = 0; i < g_n; ++i) {
        if (g_tab[i].encode == encode && g_tab[i].cols == cols && g_tab[i].bucket == b) { slot = i; break; }
    }
    if (slot == g_n) {
        if (g_n < TUNE_MAX) {
            g_n++;
        } else {
            slot = 0;
            for (uint32_t i = 1; i < g_n; ++i) { if (g_tab[i].used < g_tab[slot].used) { slot = i; } }
        }
    }
    tune_ent* e = &g_tab[slot];
    e->encode = encode; e->cols = cols; e->bucket = b; e->wpg = wpg; e->mode = mode; e->ms = ms;
    e->used = ++g_clock;
    save_locked();
    pthread_mutex_unlock(&g_mu);
}

void tc_tune_entropy(int encode, uint32_t cols, uint32_t count, uint32_t nframes, uint32_t* wpg, uint32_t* groups)
{
    (void)nframes;
    uint32_t w = 1u, mode = 0u;
    t_last_cols = cols;
    t_last_count = count;
    if (t_force == 2) {
        /* explicit shape (tcx_tune_force) */
        *wpg = t_wpg;
        *groups = t_groups;
        return;
    }
    if (t_force) {
        w = t_wpg;
        mode = t_mode;
    } else {
        pthread_mutex_lock(&g_mu);
        load_locked();
        const uint32_t b = bucket_of(count);
        for (uint32_t i = 0; i < g_n; ++i) {
            if (g_tab[i].encode == (uint32_t)(encode != 0) && g_tab[i].cols == cols && g_tab[i].bucket == b) {
                w = g_tab[i].wpg;
                mode = g_tab[i].mode;
                g_tab[i].used = ++g_clock;
                break;
            }
        }
        pthread_mutex_unlock(&g_mu);
    }
    *wpg = w;
    *groups = 0u;
    if (mode == 1u && count > 0u) {
        const uint32_t full = (count + w - 1u) / w;
        const uint32_t cap = tcg_entropy_capacity(encode, w, cols);
        if (cap > 0u && cap < full) { *groups = cap; }
    }
}

uint32_t tc_tune_tiles(int encode, uint32_t cols, uint32_t tiles)
{
    t_last_tiles = tiles;
    if (t_tpw != 0u) { return t_tpw; }
    uint32_t tpw = 1u;
    const uint32_t kind = encode ? KIND_FDCT : KIND_IDCT;
    pthread_mutex_lock(&g_mu);
    load_locked();
    const uint32_t b = bucket_of(tiles);
    for (uint32_t i = 0; i < g_n; ++i) {
        if (g_tab[i].encode == kind && g_tab[i].cols == cols && g_tab[i].bucket == b) {
            tpw = g_tab[i].wpg;
            g_tab[i].used = ++g_clock;
            break;
        }
    }
    pthread_mutex_unlock(&g_mu);
    return tpw;
}

/* ------------------------------------------------------------ tuning run */

static const uint32_t k_cand[6][2] = { { 1u, 0u }, { 2u, 0u }, { 4u, 0u }, { 1u, 1u }, { 2u, 1u }, { 4u, 1u } };

int32_t tcx_autotune_batch(const topos_frame_config* cfgs, uint32_t n, const uint64_t* dev_planes,
                           const size_t* dev_strides, const uint64_t* dev_out, const size_t* out_caps, uint32_t reps,
                           tcx_tune_report* rep)
{
    if (cfgs == NULL || n == 0u || dev_planes == NULL || dev_out == NULL || out_caps == NULL || rep == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "autotune arguments");
    }
    memset(rep, 0, sizeof(*rep));
    if (reps == 0u) { reps = 2u; }
    uint32_t* sizes = (uint32_t*)calloc(n, sizeof(uint32_t));
    uint32_t* sizes0 = (uint32_t*)calloc(n, sizeof(uint32_t));
    int32_t* rcs = (int32_t*)calloc(n, sizeof(int32_t));
    uint8_t** host = (uint8_t**)calloc(n, sizeof(uint8_t*));
    size_t* hsz = (size_t*)calloc(n, sizeof(size_t));
    uint64_t* planes = (uint64_t*)calloc((size_t)n * 4u, sizeof(uint64_t));
    topos_frame_output* infos = (topos_frame_output*)calloc(n, sizeof(topos_frame_output));
    int32_t rc = TC_OK;
    if (!sizes || !sizes0 || !rcs || !host || !hsz || !planes || !infos) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "autotune buffers");
    }
    tcg_set_timing(1);
    uint32_t enc_cols = 0u, enc_count = 0u, dec_cols = 0u, dec_count = 0u;
    float best_e = 0.0f, best_d = 0.0f;
    uint32_t be = 0u, bd = 0u;
    /* encode candidates */
    for (uint32_t k = 0; k < 6u && rc == TC_OK; ++k) {
        float ms = 0.0f;
        for (uint32_t r = 0; r < reps && rc == TC_OK; ++r) {
            t_force = 1;
            t_wpg = k_cand[k][0];
            t_mode = k_cand[k][1];
            rc = tcx_encode_batch_dev(cfgs, n, dev_planes, dev_strides, dev_out, out_caps, sizes, rcs, NULL);
            t_force = 0;
            if (rc != TC_OK) { break; }
            tcg_timing te, td;
            tc_engine_last_timing(&te, &td);
            ms = r == 0u || te.slice_enc_ms < ms ? te.slice_enc_ms : ms;
        }
        if (rc != TC_OK) { break; }
        if (k == 0u) {
            memcpy(sizes0, sizes, sizeof(uint32_t) * n);
            enc_cols = t_last_cols;
            enc_count = t_last_count;
            rep->enc_ms_default = ms;
        } else if (memcmp(sizes0, sizes, sizeof(uint32_t) * n) != 0) {
            rc = tc_fail(TC_ERR_STATE, "launch shape %u changed the packets", k);
            break;
        }
        if (k == 0u || ms < best_e) { best_e = ms; be = k; }
        rep->candidates++;
    }
    /* decode candidates on the packets just made */
    if (rc == TC_OK) {
        for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
            hsz[i] = sizes0[i];
            host[i] = (uint8_t*)malloc(hsz[i] ? hsz[i] : 1u);
            if (host[i] == NULL) { rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "autotune packets"); break; }
            rc = tcx_d2h(host[i], dev_out[i], hsz[i]);
        }
        for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
            tc_fhdr fh;
            tc_cfg_to_fhdr(&cfgs[i], &fh);
            (void)tc_fhdr_geometry(&fh);
            for (uint32_t p = 0; p < fh.plane_count && rc == TC_OK; ++p) {
                rc = tcx_dev_alloc(&planes[(size_t)i * 4u + p], (size_t)fh.pvw[p] * fh.pvh[p] * 2u);
            }
        }
    }
    for (uint32_t k = 0; k < 6u && rc == TC_OK; ++k) {
        float ms = 0.0f;
        for (uint32_t r = 0; r < reps && rc == TC_OK; ++r) {
            t_force = 1;
            t_wpg = k_cand[k][0];
            t_mode = k_cand[k][1];
            rc = tcx_decode_batch_dev(n, (const uint8_t* const*)host, hsz, dev_out, planes, NULL, infos, rcs);
            t_force = 0;
            if (rc != TC_OK) { break; }
            tcg_timing te, td;
            tc_engine_last_timing(&te, &td);
            ms = r == 0u || td.slice_dec_ms < ms ? td.slice_dec_ms : ms;
        }
        if (rc != TC_OK) { break; }
        if (k == 0u) {
            dec_cols = t_last_cols;
            dec_count = t_last_count;
            rep->dec_ms_default = ms;
        }
        if (k == 0u || ms < best_d) { best_d = ms; bd = k; }
    }
    /* transform tiles per workgroup, with the entropy shapes just chosen */
    static const uint32_t k_tpw[3] = { 1u, 2u, 4u };
    float best_f = 0.0f, best_i = 0.0f;
    uint32_t bf = 0u, bi = 0u, fq_tiles = 0u, iq_tiles = 0u;
    for (uint32_t k = 0; k < 3u && rc == TC_OK; ++k) {
        float fms = 0.0f, ims = 0.0f;
        for (uint32_t r = 0; r < reps && rc == TC_OK; ++r) {
            t_tpw = k_tpw[k];
            t_force = 1;
            t_wpg = k_cand[be][0];
            t_mode = k_cand[be][1];
            rc = tcx_encode_batch_dev(cfgs, n, dev_planes, dev_strides, dev_out, out_caps, sizes, rcs, NULL);
            fq_tiles = t_last_tiles;
            tcg_timing te, td;
            tc_engine_last_timing(&te, &td);
            if (rc == TC_OK && memcmp(sizes0, sizes, sizeof(uint32_t) * n) != 0) {
                rc = tc_fail(TC_ERR_STATE, "transform tiling %u changed the packets", k_tpw[k]);
            }
            fms = r == 0u || te.fdct_ms < fms ? te.fdct_ms : fms;
            t_wpg = k_cand[bd][0];
            t_mode = k_cand[bd][1];
            if (rc == TC_OK) {
                rc = tcx_decode_batch_dev(n, (const uint8_t* const*)host, hsz, dev_out, planes, NULL, infos, rcs);
            }
            iq_tiles = t_last_tiles;
            tc_engine_last_timing(&te, &td);
            ims = r == 0u || td.idct_ms < ims ? td.idct_ms : ims;
            t_force = 0;
            t_tpw = 0u;
        }
        if (rc != TC_OK) { break; }
        if (k == 0u) { rep->fdct_ms_default = fms; rep->idct_ms_default = ims; }
        if (k == 0u || fms < best_f) { best_f = fms; bf = k; }
        if (k == 0u || ims < best_i) { best_i = ims; bi = k; }
    }
    t_force = 0;
    t_tpw = 0u;
    tcg_set_timing(0);
    if (rc == TC_OK) {
        record(KIND_FDCT, enc_cols, fq_tiles, k_tpw[bf], 0u, best_f);
        record(KIND_IDCT, dec_cols, iq_tiles, k_tpw[bi], 0u, best_i);
        rep->fdct_tpw = k_tpw[bf];
        rep->idct_tpw = k_tpw[bi];
        rep->fdct_ms_best = best_f;
        rep->idct_ms_best = best_i;
        record(1u, enc_cols, enc_count, k_cand[be][0], k_cand[be][1], best_e);
        record(0u, dec_cols, dec_count, k_cand[bd][0], k_cand[bd][1], best_d);
        rep->enc_ms_best = best_e;
        rep->dec_ms_best = best_d;
        rep->enc_wpg = k_cand[be][0];
        rep->enc_persistent = k_cand[be][1];
        rep->dec_wpg = k_cand[bd][0];
        rep->dec_persistent = k_cand[bd][1];
    }
    for (uint32_t i = 0; planes && i < (size_t)n * 4u; ++i) { if (planes[i]) { (void)tcx_dev_free(planes[i]); } }
    for (uint32_t i = 0; host && i < n; ++i) { free(host[i]); }
    free(sizes); free(sizes0); free(rcs); free(host); free(hsz); free(planes); free(infos);
    return rc;
}

void tcx_tune_reset(void)
{
    pthread_mutex_lock(&g_mu);
    g_n = 0;
    g_loaded = 1;
    pthread_mutex_unlock(&g_mu);
}

void tcx_tune_force_tiles(uint32_t tpw)
{
    t_tpw = tpw == 1u || tpw == 2u || tpw == 4u ? tpw : 0u;
}

void tcx_tune_force(uint32_t wpg, uint32_t groups)
{
    if (wpg == 0u) {
        t_force = 0;
        return;
    }
    t_force = 2;
    t_wpg = wpg == 2u || wpg == 4u ? wpg : 1u;
    t_groups = groups;
}
