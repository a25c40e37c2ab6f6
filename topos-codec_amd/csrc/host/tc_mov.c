/* MOV container for the video track (reference native/include/topos_codec.h:
 * 670-1014, native/src/mov/mov.c): streaming mux (ftyp + mdat + moov, the
 * mdat size patched at finish), faststart rewrite (moov before mdat), and
 * the demuxer the batch decoder reads from (sample tables expanded once;
 * single packets, contiguous batches of packets, pts / duration / sync).
 * Byte layout of every atom the reference's mux writes for a video track is
 * kept (mov.c:37-46 ftyp, :386-410 tpcC, :1120-1235 stsd, :1237-1282
 * stts/stsc, :1874-2107 moov, :2109-2207 finish, :4809-5005 faststart), so
 * the same packets give the same file. Audio and timecode tracks are not
 * written (their entry points return TC_ERR_NOT_IMPLEMENTED); the reader
 * skips sound and timecode tracks. */
#define _GNU_SOURCE
#include <errno.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "tc_host.h"
#include "topos_gpu.h"

#define MOV_FTYP_BYTES 20u
#define MOV_MDAT_HDR 16u
#define MOV_MAX_SAMPLES 67108864u
#define MOV_MAX_MDAT 8

static const uint8_t MOV_FTYP[MOV_FTYP_BYTES] = { 0, 0, 0, 0x14, 'f', 't', 'y', 'p', 'q', 't', ' ', ' ',
                                                  0, 0, 2, 0, 'q', 't', ' ', ' ' };

static void be16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)(v >> 8); p[1] = (uint8_t)v; }
static void be32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}
static void be64(uint8_t* p, uint64_t v) { be32(p, (uint32_t)(v >> 32)); be32(p + 4, (uint32_t)v); }
static uint32_t ld16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }
static uint32_t ld32(const uint8_t* p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint64_t ld64(const uint8_t* p) { return ((uint64_t)ld32(p) << 32) | ld32(p + 4); }

/* ---------------------------------------------------------- atom builder */

typedef struct abuf { uint8_t* d; size_t n, cap; int err; } abuf;

static void ab_put(abuf* b, const void* p, size_t n)
{
    if (b->err) { return; }
    if (b->n + n > b->cap) {
        size_t cap = b->cap ? b->cap : 1024u;
        while (cap < b->n + n) { cap *= 2u; }
        uint8_t* d = (uint8_t*)realloc(b->d, cap);
        if (d == NULL) { b->err = 1; return; }
        b->d = d;
        b->cap = cap;
    }
    memcpy(b->d + b->n, p, n);
    b->n += n;
}
static void ab32(abuf* b, uint32_t v) { uint8_t t[4]; be32(t, v); ab_put(b, t, 4u); }
static void ab16(abuf* b, uint32_t v) { uint8_t t[2]; be16(t, v); ab_put(b, t, 2u); }
static void ab64(abuf* b, uint64_t v) { uint8_t t[8]; be64(t, v); ab_put(b, t, 8u); }
static void ab32n(abuf* b, uint32_t v, uint32_t n) { for (uint32_t i = 0; i < n; ++i) { ab32(b, v); } }
static size_t ab_open(abuf* b, const char type[4])
{
    const size_t mark = b->n;
    ab32(b, 0u);
    ab_put(b, type, 4u);
    return mark;
}
static void ab_close(abuf* b, size_t mark)
{
    if (!b->err) { be32(b->d + mark, (uint32_t)(b->n - mark)); }
}
static void ab_identity(abuf* b)
{
    for (int i = 0; i < 9; ++i) { ab32(b, (i == 0 || i == 4 || i == 8) ? 0x00010000u : 0u); }
}

/* optional sample-entry metadata carried from mux to file and back */
typedef struct mov_extra {
    const topos_alpha_budget_info* budget;  /* tpcB, alpha movies only */
    const topos_movie_meta* meta;           /* tpcD: tier, vendor, label */
    const topos_raw_meta* rmeta;            /* trwm: opaque payload */
} mov_extra;

#define TPCD_VENDOR_MAX 15u
#define TPCD_LABEL_MAX 63u
#define TPCD_PAYLOAD_MAX (12u + TPCD_VENDOR_MAX + TPCD_LABEL_MAX + 4u)

/* 24-byte tpcB: 'TPCB' ver alpha_mode alpha_bd target actual max_err flags
 * frames crc (mov.c:412-430) */
static void tpcb_build(const topos_movie_config* c, const topos_alpha_budget_info* bi, uint8_t out[24])
{
    memset(out, 0, 24u);
    memcpy(out, "TPCB", 4u);
    be16(out + 4, 1u);
    out[6] = c->alpha_mode;
    out[7] = c->alpha_bit_depth;
    be16(out + 8, bi->target_ratio_bp);
    be16(out + 10, bi->actual_ratio_bp);
    be16(out + 12, bi->max_abs_error);
    be16(out + 14, bi->flags);
    be32(out + 16, bi->frame_count);
    be32(out + 20, tc_crc32(out, 20u));
}

/* tpcD: 'TPCD' ver tier 0 vendor_len label_len vendor label crc (mov.c:432-457) */
static size_t tpcd_build(const topos_movie_meta* m, uint8_t out[TPCD_PAYLOAD_MAX])
{
    const size_t vl = strlen(m->vendor), ll = strlen(m->label);
    memset(out, 0, TPCD_PAYLOAD_MAX);
    memcpy(out, "TPCD", 4u);
    be16(out + 4, 1u);
    out[6] = m->tier_id;
    be16(out + 8, (uint32_t)vl);
    be16(out + 10, (uint32_t)ll);
    memcpy(out + 12, m->vendor, vl);
    memcpy(out + 12 + vl, m->label, ll);
    const size_t n = 12u + vl + ll;
    be32(out + n, tc_crc32(out, n));
    return n + 4u;
}

/* 28-byte tpcC payload: version = the track's bitstream major (mov.c:386-410) */
static void tpcc_build(const topos_movie_config* c, uint32_t major, uint8_t out[28])
{
    memset(out, 0, 28u);
    memcpy(out, "TPCC", 4u);
    be16(out + 4, major != 0u ? major : 1u);
    out[6] = c->profile; out[7] = c->pixel_format; out[8] = c->bit_depth; out[9] = c->qmatrix_id;
    out[10] = c->qp_base; out[11] = c->alpha_mode; out[12] = c->alpha_bit_depth;
    out[13] = (uint8_t)((c->alpha_premultiplied ? 1u : 0u) | (c->color_range ? 2u : 0u));
    out[14] = c->color_primaries; out[15] = c->color_transfer; out[16] = c->color_matrix;
    out[17] = c->chroma_siting;
    be16(out + 18, c->sar_num);
    be16(out + 20, c->sar_den);
    be32(out + 24, tc_crc32(out, 24u));
}

static void stsd_build(abuf* b, const topos_movie_config* c, uint32_t major, const mov_extra* x)
{
    static const uint8_t zero6[6] = { 0 };
    const size_t a = ab_open(b, "stsd");
    ab32(b, 0u);
    ab32(b, 1u);
    const size_t ent = ab_open(b, "TPIC");
    ab_put(b, zero6, 6u);
    ab16(b, 1u);
    ab16(b, 0u);
    ab16(b, 0u);
    ab32n(b, 0u, 3u);
    ab16(b, c->visible_width);
    ab16(b, c->visible_height);
    ab32(b, 0x00480000u);
    ab32(b, 0x00480000u);
    ab32(b, 0u);
    ab16(b, 1u);
    {
        static const char name[] = "Topos Video Codec 2.0";
        uint8_t cn[32];
        memset(cn, 0, sizeof(cn));
        cn[0] = (uint8_t)(sizeof(name) - 1u);
        memcpy(cn + 1, name, sizeof(name) - 1u);
        ab_put(b, cn, 32u);
    }
    ab16(b, 0x0018u);
    ab16(b, 0xFFFFu);
    {
        const size_t t = ab_open(b, "colr");
        ab_put(b, "nclc", 4u);
        ab16(b, c->color_primaries);
        ab16(b, c->color_transfer);
        ab16(b, c->color_matrix);
        ab_close(b, t);
    }
    if (c->sar_num != 1u || c->sar_den != 1u) {
        const size_t t = ab_open(b, "pasp");
        ab32(b, c->sar_num);
        ab32(b, c->sar_den);
        ab_close(b, t);
    }
    {
        uint8_t tp[28];
        tpcc_build(c, major, tp);
        const size_t t = ab_open(b, "tpcC");
        ab_put(b, tp, 28u);
        ab_close(b, t);
    }
    if (x->budget != NULL) {
        uint8_t tb[24];
        tpcb_build(c, x->budget, tb);
        const size_t t = ab_open(b, "tpcB");
        ab_put(b, tb, 24u);
        ab_close(b, t);
    }
    if (x->meta != NULL) {
        uint8_t td[TPCD_PAYLOAD_MAX];
        const size_t n = tpcd_build(x->meta, td);
        const size_t t = ab_open(b, "tpcD");
        ab_put(b, td, n);
        ab_close(b, t);
    }
    if (x->rmeta != NULL) {
        const size_t t = ab_open(b, "trwm");
        ab_put(b, x->rmeta->payload, x->rmeta->payload_size);
        ab_close(b, t);
    }
    ab_close(b, ent);
    ab_close(b, a);
}

/* moov of the video track (mov.c:1874-2107) */
static int32_t moov_build(abuf* b, const topos_movie_config* c, uint32_t major, uint32_t n, const uint64_t* offsets,
                          const uint32_t* sizes, const uint32_t* durs, const uint8_t* sync, uint64_t total_dur,
                          int force_co64, const mov_extra* x)
{
    int co64 = force_co64;
    for (uint32_t i = 0; i < n && !co64; ++i) { co64 = offsets[i] > 0xFFFFFFFFull; }
    const uint32_t dur32 = total_dur > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)total_dur;
    const size_t moov = ab_open(b, "moov");
    {
        const size_t t = ab_open(b, "mvhd");
        ab32n(b, 0u, 3u);
        ab32(b, 1000u);
        uint64_t ms = total_dur > UINT64_MAX / 1000u ? UINT64_MAX : total_dur * 1000u;
        ms /= c->timescale;
        ab32(b, ms > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)ms);
        ab32(b, 0x00010000u);
        ab16(b, 0x0100u);
        ab16(b, 0u);
        ab32n(b, 0u, 2u);
        ab_identity(b);
        ab32n(b, 0u, 6u);
        ab32(b, 2u);    /* next track id: the video track is 1 */
        ab_close(b, t);
    }
    const size_t trak = ab_open(b, "trak");
    {
        const size_t t = ab_open(b, "tkhd");
        ab32(b, 0x00000007u);
        ab32n(b, 0u, 2u);
        ab32(b, 1u);
        ab32(b, 0u);
        ab32(b, dur32);
        ab32n(b, 0u, 2u);
        ab16(b, 0u); ab16(b, 0u); ab16(b, 0u); ab16(b, 0u);
        ab_identity(b);
        ab32(b, (uint32_t)c->visible_width << 16);
        ab32(b, (uint32_t)c->visible_height << 16);
        ab_close(b, t);
    }
    const size_t mdia = ab_open(b, "mdia");
    {
        const size_t t = ab_open(b, "mdhd");
        ab32n(b, 0u, 3u);
        ab32(b, c->timescale);
        ab32(b, dur32);
        ab16(b, 0x55C4u);
        ab16(b, 0u);
        ab_close(b, t);
    }
    {
        static const char nm[] = "Topos Video Handler";
        const size_t t = ab_open(b, "hdlr");
        ab32n(b, 0u, 2u);
        ab_put(b, "vide", 4u);
        ab32n(b, 0u, 3u);
        ab_put(b, nm, sizeof(nm));
        ab_close(b, t);
    }
    const size_t minf = ab_open(b, "minf");
    {
        const size_t t = ab_open(b, "vmhd");
        ab32(b, 1u);
        ab16(b, 0u); ab16(b, 0u); ab16(b, 0u);
        ab_close(b, t);
    }
    {
        static const char nm[] = "Data Handler";
        const size_t t = ab_open(b, "hdlr");
        ab32n(b, 0u, 2u);
        ab_put(b, "dhlr", 4u);
        ab32n(b, 0u, 3u);
        ab_put(b, nm, sizeof(nm));
        ab_close(b, t);
    }
    {
        const size_t t = ab_open(b, "dinf");
        const size_t d = ab_open(b, "dref");
        ab32(b, 0u);
        ab32(b, 1u);
        const size_t u = ab_open(b, "url ");
        ab32(b, 1u);
        ab_close(b, u);
        ab_close(b, d);
        ab_close(b, t);
    }
    const size_t stbl = ab_open(b, "stbl");
    stsd_build(b, c, major, x);
    {
        const size_t t = ab_open(b, "stts");
        ab32(b, 0u);
        uint32_t runs = 0;
        for (uint32_t i = 0; i < n; ++i) { runs += (i == 0u || durs[i] != durs[i - 1u]) ? 1u : 0u; }
        ab32(b, runs);
        for (uint32_t i = 0; i < n;) {
            uint32_t j = i;
            while (j < n && durs[j] == durs[i]) { ++j; }
            ab32(b, j - i);
            ab32(b, durs[i]);
            i = j;
        }
        ab_close(b, t);
    }
    if (n != 0u) {
        uint32_t ns = 0;
        for (uint32_t k = 0; k < n; ++k) { ns += sync[k] ? 1u : 0u; }
        if (ns != 0u) {
            const size_t t = ab_open(b, "stss");
            ab32(b, 0u);
            ab32(b, ns);
            for (uint32_t k = 0; k < n; ++k) {
                if (sync[k]) { ab32(b, k + 1u); }
            }
            ab_close(b, t);
        }
        const size_t t = ab_open(b, "stsc");
        ab32(b, 0u); ab32(b, 1u); ab32(b, 1u); ab32(b, 1u); ab32(b, 1u);
        ab_close(b, t);
    }
    {
        const size_t t = ab_open(b, "stsz");
        ab32(b, 0u);
        ab32(b, 0u);
        ab32(b, n);
        for (uint32_t k = 0; k < n; ++k) { ab32(b, sizes[k]); }
        ab_close(b, t);
    }
    {
        const size_t t = ab_open(b, co64 ? "co64" : "stco");
        ab32(b, 0u);
        ab32(b, n);
        for (uint32_t k = 0; k < n; ++k) {
            if (co64) { ab64(b, offsets[k]); } else { ab32(b, (uint32_t)offsets[k]); }
        }
        ab_close(b, t);
    }
    ab_close(b, stbl);
    ab_close(b, minf);
    ab_close(b, mdia);
    ab_close(b, trak);
    ab_close(b, moov);
    return b->err ? tc_fail(TC_ERR_OUT_OF_MEMORY, "moov build") : TC_OK;
}

/* ----------------------------------------------------------------- mux */

struct topos_mux {
    topos_movie_config cfg;
    topos_io sink;
    uint64_t written, mdat_data_off, expect_pts;
    uint8_t finished, poisoned, major;
    uint32_t n, cap;
    uint64_t* offsets;
    uint32_t* sizes;
    uint32_t* durs;
    uint8_t* sync;
    uint8_t gop_open;
    uint16_t gop_id;
    uint8_t budget_set, meta_set, rmeta_set;
    topos_alpha_budget_info budget;
    topos_movie_meta meta;
    topos_raw_meta rmeta;
};

static mov_extra mux_extra(const topos_mux* m)
{
    mov_extra x = { m->budget_set ? &m->budget : NULL, m->meta_set ? &m->meta : NULL,
                    m->rmeta_set ? &m->rmeta : NULL };
    return x;
}

/* movie config -> frame config of the validation (mov.c:240-274) */
static void movie_to_frame(const topos_movie_config* mc, topos_frame_config* fc)
{
    memset(fc, 0, sizeof(*fc));
    fc->struct_size = (uint32_t)sizeof(*fc);
    fc->abi_version = TOPOS_CODEC_ABI_VERSION;
    fc->visible_width = mc->visible_width;
    fc->visible_height = mc->visible_height;
    fc->profile = mc->profile;
    fc->pixel_format = mc->pixel_format;
    fc->bit_depth = mc->bit_depth;
    if (mc->bit_depth == 16u) { fc->reserved[0] = 8u; }
    fc->qmatrix_id = mc->qmatrix_id;
    fc->qp_base = mc->qp_base;
    fc->qp_delta_luma = mc->qp_delta_luma;
    fc->qp_delta_chroma = mc->qp_delta_chroma;
    fc->slice_rows = 16u;
    fc->alpha_mode = mc->alpha_mode;
    fc->alpha_bit_depth = mc->alpha_bit_depth;
    fc->alpha_premultiplied = mc->alpha_premultiplied;
    fc->color_range = mc->color_range;
    fc->color_primaries = mc->color_primaries;
    fc->color_transfer = mc->color_transfer;
    fc->color_matrix = mc->color_matrix;
    fc->chroma_siting = mc->chroma_siting;
    fc->sar_num = mc->sar_num;
    fc->sar_den = mc->sar_den;
}

static int32_t sink_write(topos_mux* m, const void* p, size_t n)
{
    if (n == 0u) { return TC_OK; }
    if (m->sink.write(m->sink.ctx, p, n) != TC_OK) { return tc_fail(TC_ERR_IO, "sink write failed"); }
    m->written += n;
    return TC_OK;
}

int32_t tc_mux_create(const topos_movie_config* cfg, const topos_io* sink, topos_mux** out)
{
    if (sink == NULL || out == NULL || sink->struct_size != (uint32_t)sizeof(topos_io) ||
        sink->abi_version != TOPOS_CODEC_ABI_VERSION || sink->write == NULL || sink->seek_write == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "mux sink needs write + seek_write");
    }
    if (cfg == NULL || cfg->struct_size != (uint32_t)sizeof(topos_movie_config) ||
        cfg->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "movie config NULL or struct_size/abi_version mismatch");
    }
    topos_movie_config nc = *cfg;
    if (nc.timescale == 0u) { nc.timescale = 24000u; }
    if (nc.timescale > 1000000u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "timescale outside [1, 1000000]"); }
    /* reserved[0..4]: the sound track declaration (codec 0 = none) */
    if (nc.reserved[0] != 0u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "audio tracks are not written by this backend");
    }
    if ((nc.reserved[1] | nc.reserved[2] | nc.reserved[3] | nc.reserved[4]) != 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "audio codec none needs the other audio slots at 0");
    }
    topos_mux* m = (topos_mux*)calloc(1, sizeof(topos_mux));
    if (m == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "mux alloc"); }
    m->cfg = nc;
    topos_frame_config fc;
    movie_to_frame(&m->cfg, &fc);
    int32_t rc = tc_frame_config_validate(&fc);
    if (rc != TC_OK) { free(m); return rc; }
    m->sink = *sink;
    uint8_t mdat[MOV_MDAT_HDR];
    be32(mdat, 1u);
    memcpy(mdat + 4, "mdat", 4u);
    be64(mdat + 8, 0u);
    rc = sink_write(m, MOV_FTYP, MOV_FTYP_BYTES);
    if (rc == TC_OK) { rc = sink_write(m, mdat, MOV_MDAT_HDR); }
    if (rc != TC_OK) { free(m); return rc; }
    m->mdat_data_off = m->written;
    *out = m;
    return TC_OK;
}

static int32_t mux_grow(topos_mux* m)
{
    if (m->n < m->cap) { return TC_OK; }
    if (m->cap >= MOV_MAX_SAMPLES) { return tc_fail(TC_ERR_LIMIT_EXCEEDED, "sample count"); }
    const uint32_t cap = m->cap ? m->cap * 2u : 64u;
    uint64_t* o = (uint64_t*)realloc(m->offsets, (size_t)cap * sizeof(uint64_t));
    if (o) { m->offsets = o; }
    uint32_t* s = (uint32_t*)realloc(m->sizes, (size_t)cap * sizeof(uint32_t));
    if (s) { m->sizes = s; }
    uint32_t* d = (uint32_t*)realloc(m->durs, (size_t)cap * sizeof(uint32_t));
    if (d) { m->durs = d; }
    uint8_t* y = (uint8_t*)realloc(m->sync, cap);
    if (y) { m->sync = y; }
    if (!o || !s || !d || !y) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "mux index grow"); }
    m->cap = cap;
    return TC_OK;
}

/* packet header fields that must match the movie configuration (mov.c:644-657) */
static int cfg_mismatch(const tc_fhdr* fh, const topos_movie_config* c, int sar_default)
{
    uint32_t sn = fh->sar_num, sd = fh->sar_den;
    if (sar_default && sn == 0u && sd == 0u) { sn = 1u; sd = 1u; }
    return fh->visible_w != c->visible_width || fh->visible_h != c->visible_height || fh->profile != c->profile ||
           fh->pixel_format != c->pixel_format || fh->bit_depth != c->bit_depth ||
           fh->qmatrix_id != c->qmatrix_id || fh->alpha_mode != c->alpha_mode ||
           fh->alpha_bit_depth != c->alpha_bit_depth ||
           ((fh->flags & 1u) != 0u) != (c->alpha_premultiplied != 0u) || fh->color_range != c->color_range ||
           fh->color_primaries != c->color_primaries || fh->color_transfer != c->color_transfer ||
           fh->color_matrix != c->color_matrix || fh->chroma_siting != c->chroma_siting || sn != c->sar_num ||
           sd != c->sar_den;
}

int32_t tc_mux_add_packet(topos_mux* m, const uint8_t* packet, size_t size, uint64_t pts_tick, uint32_t dur_tick)
{
    if (m == NULL || m->finished || m->poisoned) {
        return tc_fail(TC_ERR_STATE, "mux not created, finished or failed (free only)");
    }
    if (packet == NULL || size == 0u || dur_tick == 0u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet/dur"); }
    if (pts_tick != m->expect_pts) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "pts must equal the previous pts + dur (no edit list)");
    }
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    int32_t rc = tc_parse_structure(packet, size, v);
    const tc_fhdr fh = v->fh;
    free(v);
    if (rc != TC_OK) { return rc; }
    if (m->major != 0u && m->major != fh.major) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "mixed bitstream major (track %u, packet %u)", (unsigned)m->major,
                       (unsigned)fh.major);
    }
    if (cfg_mismatch(&fh, &m->cfg, 0)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet header differs from the movie configuration (tpcC rule)");
    }
    uint8_t sync = 1u;
    if (fh.major == 9u || (fh.major == 7u && fh.entropy_mode == 8u)) {
        if (fh.frame_type == 0u) {
            m->gop_open = 1u;
            m->gop_id = fh.gop_id;
        } else {
            if (!m->gop_open) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "first sample of a GOP chain must be I"); }
            if (fh.gop_id != m->gop_id) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "P sample of another GOP"); }
            sync = 0u;
        }
    }
    if ((rc = mux_grow(m)) != TC_OK) { return rc; }
    const uint64_t off = m->written;
    if ((rc = sink_write(m, packet, size)) != TC_OK) {
        m->poisoned = 1u;
        return rc;
    }
    m->offsets[m->n] = off;
    m->sizes[m->n] = (uint32_t)size;
    m->durs[m->n] = dur_tick;
    m->sync[m->n] = sync;
    m->n++;
    m->expect_pts += dur_tick;
    if (m->major == 0u) { m->major = fh.major; }
    return TC_OK;
}

int32_t tc_mux_finish(topos_mux* m)
{
    if (m == NULL || m->finished || m->poisoned) {
        return tc_fail(TC_ERR_STATE, "mux not created, finished or failed (free only)");
    }
    uint64_t total = 0;
    for (uint32_t i = 0; i < m->n; ++i) { total += m->durs[i]; }
    abuf b = { NULL, 0, 0, 0 };
    const mov_extra x = mux_extra(m);
    int32_t rc = moov_build(&b, &m->cfg, m->major, m->n, m->offsets, m->sizes, m->durs, m->sync, total, 0, &x);
    const uint64_t payload_end = m->written;
    if (rc == TC_OK) { rc = sink_write(m, b.d, b.n); }
    if (rc == TC_OK) {
        uint8_t sz[8];
        be64(sz, payload_end - m->mdat_data_off + MOV_MDAT_HDR);
        if (m->sink.seek_write(m->sink.ctx, (uint64_t)MOV_FTYP_BYTES + 8u, sz, 8u) != TC_OK) {
            rc = tc_fail(TC_ERR_IO, "mdat size patch failed");
        }
    }
    free(b.d);
    m->finished = 1u;
    return rc;
}

/* tpcB metadata, alpha movies only (mov.c:712-738) */
int32_t tc_mux_set_alpha_budget(topos_mux* m, const topos_alpha_budget_info* info)
{
    if (m == NULL || m->finished || m->poisoned) {
        return tc_fail(TC_ERR_STATE, "mux not created, finished or failed (free only)");
    }
    if (info == NULL || info->struct_size != (uint32_t)sizeof(topos_alpha_budget_info) ||
        info->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "budget info NULL or struct mismatch");
    }
    if (m->cfg.alpha_mode == 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "a movie without alpha carries no alpha budget");
    }
    if (info->target_ratio_bp > 10000u && info->target_ratio_bp != TC_ALPHA_BUDGET_RATIO_UNSET) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "target ratio above 10000 bp");
    }
    m->budget = *info;
    m->budget_set = 1u;
    return TC_OK;
}

/* tpcD metadata: tier, vendor and label (mov.c:740-772) */
int32_t tc_mux_set_movie_meta(topos_mux* m, const topos_movie_meta* meta)
{
    if (m == NULL || m->finished || m->poisoned) {
        return tc_fail(TC_ERR_STATE, "mux not created, finished or failed (free only)");
    }
    if (meta == NULL || meta->struct_size != (uint32_t)sizeof(topos_movie_meta) ||
        meta->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "movie meta NULL or struct mismatch");
    }
    if (meta->tier_id > TC_TIER_RAW) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "tier_id outside 0..8"); }
    if (meta->reserved[0] != 0u || meta->reserved[1] != 0u || meta->reserved[2] != 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "movie meta reserved bytes must be 0");
    }
    if (strnlen(meta->vendor, sizeof(meta->vendor)) >= sizeof(meta->vendor)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "vendor longer than 15 bytes");
    }
    if (strnlen(meta->label, sizeof(meta->label)) >= sizeof(meta->label)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "label longer than 63 bytes");
    }
    m->meta = *meta;
    m->meta_set = 1u;
    return TC_OK;
}

/* trwm payload, carried opaque (mov.c:776-797) */
int32_t tc_mux_set_raw_meta(topos_mux* m, const topos_raw_meta* meta)
{
    if (m == NULL || meta == NULL || meta->struct_size != (uint32_t)sizeof(topos_raw_meta) ||
        meta->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "set_raw_meta: NULL or struct mismatch");
    }
    if (meta->reserved != 0u || meta->payload_size == 0u || meta->payload_size > TC_RAW_META_MAX) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "raw meta payload size outside 1..256 or reserved set");
    }
    if (m->finished || m->poisoned) {
        return tc_fail(TC_ERR_STATE, "mux finished or failed (free only)");
    }
    m->rmeta = *meta;
    m->rmeta_set = 1u;
    return TC_OK;
}

/* sound and timecode tracks are outside this backend (reference
 * topos_codec.h:850-874, 938-945) */
int32_t tc_mux_add_audio(topos_mux* m, const uint8_t* data, size_t size, uint32_t num_samples)
{
    (void)m; (void)data; (void)size; (void)num_samples;
    return tc_fail(TC_ERR_NOT_IMPLEMENTED, "audio tracks are not written by this backend");
}
int32_t tc_mux_set_audio_asc(topos_mux* m, const uint8_t* asc, size_t size)
{
    (void)m; (void)asc; (void)size;
    return tc_fail(TC_ERR_NOT_IMPLEMENTED, "audio tracks are not written by this backend");
}
int32_t tc_mux_set_timecode(topos_mux* m, uint32_t hh, uint32_t mm, uint32_t ss, uint32_t ff, uint32_t fps,
                            uint32_t drop_frame)
{
    (void)m; (void)hh; (void)mm; (void)ss; (void)ff; (void)fps; (void)drop_frame;
    return tc_fail(TC_ERR_NOT_IMPLEMENTED, "timecode tracks are not written by this backend");
}
int32_t tc_mux_set_audio_priming(topos_mux* m, uint32_t samples)
{
    (void)m; (void)samples;
    return tc_fail(TC_ERR_NOT_IMPLEMENTED, "audio tracks are not written by this backend");
}

void tc_mux_free(topos_mux* m)
{
    if (m == NULL) { return; }
    free(m->offsets);
    free(m->sizes);
    free(m->durs);
    free(m->sync);
    free(m);
}

/* -------------------------------------------------------------- demux */

struct topos_movie {
    topos_io io;
    topos_movie_config cfg;
    uint32_t width, height, timescale, n, faststart, major;
    uint64_t* offsets;
    uint64_t* pts;
    uint32_t* sizes;
    uint32_t* durs;
    uint8_t* sync;          /* bit per sample */
    uint32_t index_bytes;
    uint8_t budget_set, meta_set, rmeta_set;
    topos_alpha_budget_info budget;
    topos_movie_meta meta;
    topos_raw_meta rmeta;
    int fd;                 /* tc_movie_open_fd: pread source, else -1 */
};

typedef struct atom { uint64_t size, body; char type[4]; int ext; } atom;

static int32_t io_read(const topos_movie* mv, uint64_t off, void* buf, size_t len)
{
    if (len == 0u) { return TC_OK; }
    if (off > mv->io.length || len > mv->io.length - off) { return tc_fail(TC_ERR_TRUNCATED, "read past end"); }
    if (mv->io.read(mv->io.ctx, off, buf, len) != TC_OK) { return tc_fail(TC_ERR_IO, "io read failed"); }
    return TC_OK;
}

/* atom header at off inside [off, end) (mov.c:2305-2346) */
static int32_t read_atom(const topos_movie* mv, uint64_t off, uint64_t end, atom* h)
{
    if (off > end || end - off < 8u) { return tc_fail(TC_ERR_MALFORMED, "atom header out of range"); }
    uint8_t hd[8];
    int32_t rc = io_read(mv, off, hd, 8u);
    if (rc != TC_OK) { return rc; }
    h->size = ld32(hd);
    memcpy(h->type, hd + 4, 4u);
    h->ext = 0;
    h->body = off + 8u;
    if (h->size == 1u) {
        if (end - off < 16u) { return tc_fail(TC_ERR_MALFORMED, "extended size header out of range"); }
        uint8_t e[8];
        if ((rc = io_read(mv, h->body, e, 8u)) != TC_OK) { return rc; }
        h->size = ld64(e);
        if (h->size < 16u) { return tc_fail(TC_ERR_MALFORMED, "extended size < 16"); }
        h->ext = 1;
        h->body = off + 16u;
    } else if (h->size == 0u) {
        h->size = end - off;
    } else if (h->size < 8u) {
        return tc_fail(TC_ERR_MALFORMED, "atom size < 8");
    }
    if (h->size > end - off) { return tc_fail(TC_ERR_MALFORMED, "atom exceeds its parent"); }
    return TC_OK;
}

static int is(const atom* h, const char t[4]) { return memcmp(h->type, t, 4u) == 0; }
static uint64_t body_len(const atom* h) { return h->size - (h->ext ? 16u : 8u); }

typedef struct tables {
    uint64_t stts, stts_end, stsz, stsz_end, stsc, stco, stco_end, stss, stss_end;
    uint64_t total;
    uint32_t stsz_size, stsz_n, stsc_n, stco_n, stss_n;
    int co64, has_stsd, has_stts, has_stsz, has_stsc, has_stco, has_stss, has_tpcc, has_tpcb, has_tpcd;
    uint8_t tpcc[28], tpcb[24];
    uint32_t entry_w, entry_h;
} tables;

static int32_t full_count(const topos_movie* mv, uint64_t body, uint64_t end, uint32_t* count)
{
    if (body > end || end - body < 8u) { return tc_fail(TC_ERR_MALFORMED, "full atom too short"); }
    uint8_t b[8];
    int32_t rc = io_read(mv, body, b, 8u);
    if (rc == TC_OK) { *count = ld32(b + 4); }
    return rc;
}

/* TPIC sample entry: geometry, tpcC, tpcB, tpcD, trwm (mov.c:2392-2546) */
static int32_t parse_stsd(topos_movie* mv, uint64_t body, uint64_t end, tables* t)
{
    uint32_t cnt = 0;
    int32_t rc = full_count(mv, body, end, &cnt);
    if (rc != TC_OK) { return rc; }
    if (cnt != 1u) { return tc_fail(TC_ERR_MALFORMED, "stsd entry_count %u != 1", cnt); }
    atom e;
    if ((rc = read_atom(mv, body + 8u, end, &e)) != TC_OK) { return rc; }
    if (!is(&e, "TPIC")) { return tc_fail(TC_ERR_MALFORMED, "sample entry is not TPIC"); }
    if (body_len(&e) < 78u) { return tc_fail(TC_ERR_MALFORMED, "TPIC entry too short"); }
    uint8_t f[70];
    if ((rc = io_read(mv, e.body, f, 70u)) != TC_OK) { return rc; }
    t->entry_w = ld16(f + 24);
    t->entry_h = ld16(f + 26);
    const uint64_t eend = e.body + body_len(&e);
    for (uint64_t off = e.body + 78u; off < eend;) {
        atom c;
        if ((rc = read_atom(mv, off, eend, &c)) != TC_OK) { return rc; }
        const uint64_t cl = body_len(&c);
        if (is(&c, "tpcC")) {
            if (cl != 28u) { return tc_fail(TC_ERR_MALFORMED, "tpcC length != 28"); }
            uint8_t tp[28];
            if ((rc = io_read(mv, c.body, tp, 28u)) != TC_OK) { return rc; }
            const uint32_t ver = ld16(tp + 4);
            if (memcmp(tp, "TPCC", 4u) != 0 || ver < 1u || ver > 9u) {
                return tc_fail(TC_ERR_MALFORMED, "tpcC magic/version");
            }
            mv->major = ver;
            if (tc_crc32(tp, 24u) != ld32(tp + 24)) { return tc_fail(TC_ERR_CHECKSUM_MISMATCH, "tpcC CRC"); }
            if (t->has_tpcc) { return tc_fail(TC_ERR_MALFORMED, "duplicate tpcC"); }
            memcpy(t->tpcc, tp, 28u);
            t->has_tpcc = 1;
        } else if (is(&c, "tpcB")) {
            if (cl != 24u) { return tc_fail(TC_ERR_MALFORMED, "tpcB length != 24"); }
            uint8_t tb[24];
            if ((rc = io_read(mv, c.body, tb, 24u)) != TC_OK) { return rc; }
            if (memcmp(tb, "TPCB", 4u) != 0 || ld16(tb + 4) != 1u) {
                return tc_fail(TC_ERR_MALFORMED, "tpcB magic/version");
            }
            if (tc_crc32(tb, 20u) != ld32(tb + 20)) { return tc_fail(TC_ERR_CHECKSUM_MISMATCH, "tpcB CRC"); }
            if (t->has_tpcb) { return tc_fail(TC_ERR_MALFORMED, "duplicate tpcB"); }
            memcpy(t->tpcb, tb, 24u);
            t->has_tpcb = 1;
        } else if (is(&c, "tpcD")) {
            if (cl < 16u || cl > TPCD_PAYLOAD_MAX) { return tc_fail(TC_ERR_MALFORMED, "tpcD length"); }
            uint8_t td[TPCD_PAYLOAD_MAX];
            if ((rc = io_read(mv, c.body, td, (size_t)cl)) != TC_OK) { return rc; }
            if (memcmp(td, "TPCD", 4u) != 0 || ld16(td + 4) != 1u) {
                return tc_fail(TC_ERR_MALFORMED, "tpcD magic/version");
            }
            if (td[7] != 0u || td[6] > TC_TIER_RAW) { return tc_fail(TC_ERR_MALFORMED, "tpcD reserved/tier"); }
            const uint32_t vl = ld16(td + 8), ll = ld16(td + 10);
            if (12u + (uint64_t)vl + ll + 4u != cl || vl > TPCD_VENDOR_MAX || ll > TPCD_LABEL_MAX) {
                return tc_fail(TC_ERR_MALFORMED, "tpcD string lengths");
            }
            if (tc_crc32(td, (size_t)cl - 4u) != ld32(td + cl - 4u)) {
                return tc_fail(TC_ERR_CHECKSUM_MISMATCH, "tpcD CRC");
            }
            if (t->has_tpcd) { return tc_fail(TC_ERR_MALFORMED, "duplicate tpcD"); }
            memset(&mv->meta, 0, sizeof(mv->meta));
            mv->meta.struct_size = (uint32_t)sizeof(mv->meta);
            mv->meta.abi_version = TOPOS_CODEC_ABI_VERSION;
            mv->meta.tier_id = td[6];
            memcpy(mv->meta.vendor, td + 12, vl);
            memcpy(mv->meta.label, td + 12 + vl, ll);
            t->has_tpcd = 1;
            mv->meta_set = 1u;
        } else if (is(&c, "trwm")) {
            if (cl < 8u || cl > TC_RAW_META_MAX) { return tc_fail(TC_ERR_MALFORMED, "trwm length"); }
            if (mv->rmeta_set) { return tc_fail(TC_ERR_MALFORMED, "duplicate trwm"); }
            if ((rc = io_read(mv, c.body, mv->rmeta.payload, (size_t)cl)) != TC_OK) { return rc; }
            mv->rmeta.struct_size = (uint32_t)sizeof(mv->rmeta);
            mv->rmeta.abi_version = TOPOS_CODEC_ABI_VERSION;
            mv->rmeta.payload_size = (uint16_t)cl;
            mv->rmeta.reserved = 0u;
            mv->rmeta_set = 1u;
        }
        off += c.size;
    }
    return t->has_tpcc ? TC_OK : tc_fail(TC_ERR_MALFORMED, "missing tpcC");
}

/* entry count + table bounds of a full atom (mov.c:2548-2656) */
static int32_t table(topos_movie* mv, uint64_t body, uint64_t end, uint32_t entry, uint32_t* n, uint64_t* off,
                     uint64_t* tend, const char* what)
{
    int32_t rc = full_count(mv, body, end, n);
    if (rc != TC_OK) { return rc; }
    *off = body + 8u;
    *tend = *off + (uint64_t)*n * entry;
    if (*tend > end) { return tc_fail(TC_ERR_MALFORMED, "%s table exceeds its atom", what); }
    return TC_OK;
}

static int32_t parse_stts(topos_movie* mv, uint64_t body, uint64_t end, tables* t)
{
    uint32_t n = 0;
    int32_t rc = table(mv, body, end, 8u, &n, &t->stts, &t->stts_end, "stts");
    for (uint32_t i = 0; i < n && rc == TC_OK; ++i) {
        uint8_t e[8];
        if ((rc = io_read(mv, t->stts + (uint64_t)i * 8u, e, 8u)) != TC_OK) { break; }
        if (ld32(e + 4) == 0u) { return tc_fail(TC_ERR_MALFORMED, "stts delta 0"); }
        t->total += ld32(e);
    }
    return rc;
}

static int32_t parse_stbl(topos_movie* mv, uint64_t s, uint64_t send, tables* t)
{
    int32_t rc = TC_OK;
    while (s < send && rc == TC_OK) {
        atom c;
        if ((rc = read_atom(mv, s, send, &c)) != TC_OK) { return rc; }
        const uint64_t cend = s + c.size;
        int seen = 0;
        if (is(&c, "stsd")) {
            seen = t->has_stsd; t->has_stsd = 1;
            if (!seen) { rc = parse_stsd(mv, c.body, cend, t); }
        } else if (is(&c, "stts")) {
            seen = t->has_stts; t->has_stts = 1;
            if (!seen) { rc = parse_stts(mv, c.body, cend, t); }
        } else if (is(&c, "stsz")) {
            seen = t->has_stsz; t->has_stsz = 1;
            if (!seen) {
                if (c.body > cend || cend - c.body < 12u) { return tc_fail(TC_ERR_MALFORMED, "stsz too short"); }
                uint8_t b[12];
                if ((rc = io_read(mv, c.body, b, 12u)) != TC_OK) { return rc; }
                t->stsz_size = ld32(b + 4);
                t->stsz_n = ld32(b + 8);
                t->stsz = c.body + 12u;
                t->stsz_end = t->stsz + (uint64_t)t->stsz_n * 4u;
                if (t->stsz_size == 0u && t->stsz_end > cend) {
                    return tc_fail(TC_ERR_MALFORMED, "stsz table exceeds its atom");
                }
            }
        } else if (is(&c, "stsc")) {
            seen = t->has_stsc; t->has_stsc = 1;
            uint64_t e = 0;
            if (!seen) { rc = table(mv, c.body, cend, 12u, &t->stsc_n, &t->stsc, &e, "stsc"); }
        } else if (is(&c, "stco") || is(&c, "co64")) {
            seen = t->has_stco; t->has_stco = 1;
            if (!seen) {
                t->co64 = is(&c, "co64");
                rc = table(mv, c.body, cend, t->co64 ? 8u : 4u, &t->stco_n, &t->stco, &t->stco_end, "stco/co64");
            }
        } else if (is(&c, "stss")) {
            seen = t->has_stss; t->has_stss = 1;
            if (!seen) { rc = table(mv, c.body, cend, 4u, &t->stss_n, &t->stss, &t->stss_end, "stss"); }
        }
        if (seen) { return tc_fail(TC_ERR_MALFORMED, "duplicate sample table atom"); }
        s = cend;
    }
    return rc;
}

/* handler type of a trak: the hdlr inside its mdia, zeros when absent */
static int32_t trak_handler(const topos_movie* mv, uint64_t body, uint64_t end, char out[4])
{
    memset(out, 0, 4u);
    for (uint64_t o = body; o < end;) {
        atom c;
        int32_t rc = read_atom(mv, o, end, &c);
        if (rc != TC_OK) { return rc; }
        if (is(&c, "mdia")) {
            for (uint64_t d = c.body; d < o + c.size;) {
                atom h;
                if ((rc = read_atom(mv, d, o + c.size, &h)) != TC_OK) { return rc; }
                if (is(&h, "hdlr") && h.size >= 24u) { return io_read(mv, h.body + 8u, out, 4u); }
                d += h.size;
            }
        }
        o += c.size;
    }
    return TC_OK;
}

/* the video trak: mdhd timescale, hdlr, the sample tables (mov.c:3822-3910) */
static int32_t parse_video_trak(topos_movie* mv, uint64_t body, uint64_t end, tables* t, int* have_vide,
                                int* have_stbl)
{
    int32_t rc = TC_OK;
    for (uint64_t tk = body; tk < end;) {
        atom tc;
        if ((rc = read_atom(mv, tk, end, &tc)) != TC_OK) { return rc; }
        if (is(&tc, "edts")) { return tc_fail(TC_ERR_MALFORMED, "video track with edts"); }
        if (is(&tc, "mdia")) {
            const uint64_t dend = tk + tc.size;
            for (uint64_t d = tc.body; d < dend;) {
                atom m;
                if ((rc = read_atom(mv, d, dend, &m)) != TC_OK) { return rc; }
                if (is(&m, "mdhd")) {
                    if (m.size < 32u) { return tc_fail(TC_ERR_MALFORMED, "mdhd too short"); }
                    uint8_t b[4];
                    if ((rc = io_read(mv, m.body + 12u, b, 4u)) != TC_OK) { return rc; }
                    mv->timescale = ld32(b);
                    if (mv->timescale == 0u) { return tc_fail(TC_ERR_MALFORMED, "timescale 0"); }
                } else if (is(&m, "hdlr")) {
                    if (m.size >= 24u) {
                        uint8_t b[4];
                        if ((rc = io_read(mv, m.body + 8u, b, 4u)) != TC_OK) { return rc; }
                        *have_vide |= memcmp(b, "vide", 4u) == 0;
                    }
                } else if (is(&m, "minf")) {
                    const uint64_t fend = d + m.size;
                    for (uint64_t f = m.body; f < fend;) {
                        atom s;
                        if ((rc = read_atom(mv, f, fend, &s)) != TC_OK) { return rc; }
                        if (is(&s, "stbl")) {
                            *have_stbl = 1;
                            if ((rc = parse_stbl(mv, s.body, f + s.size, t)) != TC_OK) { return rc; }
                        }
                        f += s.size;
                    }
                }
                d += m.size;
            }
        }
        tk += tc.size;
    }
    return rc;
}

static void movie_free(topos_movie* mv)
{
    if (mv == NULL) { return; }
    free(mv->offsets);
    free(mv->pts);
    free(mv->sizes);
    free(mv->durs);
    free(mv->sync);
    free(mv);
}

/* stsc runs x stco chunks -> sample offsets; runs start at chunk 1, are
 * strictly increasing and contiguous, every chunk is used (mov.c:2658-2733) */
static int32_t expand_offsets(topos_movie* mv, const tables* t)
{
    uint32_t si = 0, ci = 0;
    int32_t rc = TC_OK;
    for (uint32_t r = 0; r < t->stsc_n && si < mv->n; ++r) {
        uint8_t e[12];
        if ((rc = io_read(mv, t->stsc + (uint64_t)r * 12u, e, 12u)) != TC_OK) { return rc; }
        const uint32_t first = ld32(e), spc = ld32(e + 4);
        uint64_t run = UINT64_MAX;
        if (r + 1u < t->stsc_n) {
            uint8_t nx[12];
            if ((rc = io_read(mv, t->stsc + (uint64_t)(r + 1u) * 12u, nx, 12u)) != TC_OK) { return rc; }
            const uint32_t next = ld32(nx);
            if (next <= first) { return tc_fail(TC_ERR_MALFORMED, "stsc first_chunk not increasing"); }
            run = next - first;
        }
        if (first < 1u || spc == 0u || first != ci + 1u) {
            return tc_fail(TC_ERR_MALFORMED, "stsc entry invalid or not contiguous");
        }
        for (uint64_t k = 0; k < run && si < mv->n; ++k) {
            if (ci >= t->stco_n) { return tc_fail(TC_ERR_MALFORMED, "stco has too few chunks"); }
            uint8_t o8[8];
            const uint32_t es = t->co64 ? 8u : 4u;
            if ((rc = io_read(mv, t->stco + (uint64_t)ci * es, o8, es)) != TC_OK) { return rc; }
            uint64_t off = t->co64 ? ld64(o8) : ld32(o8);
            ci++;
            for (uint32_t j = 0; j < spc && si < mv->n; ++j) {
                mv->offsets[si] = off;
                off += mv->sizes[si];
                si++;
            }
        }
    }
    if (si != mv->n || ci != t->stco_n) {
        return tc_fail(TC_ERR_MALFORMED, "stsc/stco expansion does not match the sample count");
    }
    return TC_OK;
}

int32_t tc_movie_open(const topos_io* io, topos_movie** out)
{
    if (io == NULL || out == NULL || io->read == NULL || io->struct_size != (uint32_t)sizeof(topos_io) ||
        io->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "io NULL, struct mismatch or no read");
    }
    *out = NULL;
    if (io->length < 20u) { return tc_fail(TC_ERR_TRUNCATED, "file too short"); }
    topos_movie* mv = (topos_movie*)calloc(1, sizeof(topos_movie));
    if (mv == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "movie alloc"); }
    mv->io = *io;
    mv->fd = -1;
    int32_t rc = TC_OK;
    uint64_t moov = UINT64_MAX, moov_end = 0, moov_body = 0, mdat = UINT64_MAX;
    uint64_t mlo[MOV_MAX_MDAT], mhi[MOV_MAX_MDAT];
    int nm = 0, ftyp = 0;
    uint64_t off = 0;
    /* top level; an mdat whose size runs past the end of the file (a cut
     * download) ends at the file end, its samples judged against the declared
     * end (mov.c:3690-3745) */
    while (off + 8u <= io->length) {
        atom h;
        uint64_t declared_end = 0;
        if ((rc = read_atom(mv, off, io->length, &h)) != TC_OK) {
            uint8_t h8[8];
            if (io_read(mv, off, h8, 8u) != TC_OK || memcmp(h8 + 4, "mdat", 4u) != 0) { goto fail; }
            rc = TC_OK;
            memcpy(h.type, "mdat", 4u);
            h.size = io->length - off;
            h.ext = 0;
            h.body = off + 8u;
            if (ld32(h8) == 1u) {
                uint8_t e8[8];
                if (io_read(mv, off + 8u, e8, 8u) == TC_OK) {
                    const uint64_t d = ld64(e8);
                    if (d >= 16u && d <= UINT64_MAX - off) { declared_end = off + d; }
                }
            }
            if (declared_end == 0u) { declared_end = io->length; }
        }
        if (is(&h, "ftyp")) {
            ftyp = 1;
        } else if (is(&h, "moov")) {
            if (moov != UINT64_MAX) { rc = tc_fail(TC_ERR_MALFORMED, "duplicate moov"); goto fail; }
            moov = off;
            moov_end = off + h.size;
            moov_body = h.body;
        } else if (is(&h, "mdat")) {
            mdat = off;
            if (nm >= MOV_MAX_MDAT) { rc = tc_fail(TC_ERR_MALFORMED, "more than 8 mdat"); goto fail; }
            mlo[nm] = h.body;
            mhi[nm] = declared_end ? declared_end : off + h.size;
            nm++;
        }
        off += h.size;
    }
    if (off != io->length) { rc = tc_fail(TC_ERR_MALFORMED, "top-level atoms do not close"); goto fail; }
    if (!ftyp || moov == UINT64_MAX) { rc = tc_fail(TC_ERR_MALFORMED, "missing ftyp or moov"); goto fail; }
    if (nm == 0) { rc = tc_fail(TC_ERR_MALFORMED, "missing mdat"); goto fail; }
    mv->faststart = mdat != UINT64_MAX && moov < mdat;

    tables t;
    memset(&t, 0, sizeof(t));
    {
        int have_trak = 0, have_vide = 0, have_stbl = 0;
        for (uint64_t o = moov_body; o < moov_end;) {
            atom c;
            if ((rc = read_atom(mv, o, moov_end, &c)) != TC_OK) { goto fail; }
            if (is(&c, "trak")) {
                char ht[4];
                if ((rc = trak_handler(mv, c.body, o + c.size, ht)) != TC_OK) { goto fail; }
                /* sound and timecode tracks are skipped; any other trak is the video track */
                if (memcmp(ht, "soun", 4u) != 0 && memcmp(ht, "tmcd", 4u) != 0) {
                    if (have_trak) { rc = tc_fail(TC_ERR_MALFORMED, "more than one video trak"); goto fail; }
                    have_trak = 1;
                    rc = parse_video_trak(mv, c.body, o + c.size, &t, &have_vide, &have_stbl);
                    if (rc != TC_OK) { goto fail; }
                }
            }
            o += c.size;
        }
        if (!have_trak || !have_vide || !have_stbl) {
            rc = tc_fail(TC_ERR_MALFORMED, "moov lacks trak/vide/stbl");
            goto fail;
        }
    }
    if (t.total == 0u) { rc = tc_fail(TC_ERR_MALFORMED, "empty stts"); goto fail; }
    if (t.total > MOV_MAX_SAMPLES) {
        rc = tc_fail(TC_ERR_LIMIT_EXCEEDED, "sample count %llu", (unsigned long long)t.total);
        goto fail;
    }
    if (t.stsz_n != t.total) { rc = tc_fail(TC_ERR_MALFORMED, "stsz count != stts samples"); goto fail; }
    mv->n = (uint32_t)t.total;
    mv->offsets = (uint64_t*)malloc((size_t)mv->n * sizeof(uint64_t));
    mv->pts = (uint64_t*)malloc((size_t)mv->n * sizeof(uint64_t));
    mv->sizes = (uint32_t*)malloc((size_t)mv->n * sizeof(uint32_t));
    mv->durs = (uint32_t*)malloc((size_t)mv->n * sizeof(uint32_t));
    mv->sync = (uint8_t*)calloc((size_t)mv->n / 8u + 1u, 1u);
    if (!mv->offsets || !mv->pts || !mv->sizes || !mv->durs || !mv->sync) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "sample index");
        goto fail;
    }
    /* offsets + pts (u64) and sizes + durations (u32) per sample, the sync bitmap */
    mv->index_bytes = (uint32_t)((size_t)mv->n * (sizeof(uint64_t) + 2u * sizeof(uint32_t) + sizeof(uint64_t)) +
                                 mv->n / 8u + 1u);
    {
        uint32_t si = 0;
        uint64_t acc = 0;
        for (uint64_t p = t.stts; p < t.stts_end && si < mv->n; p += 8u) {
            uint8_t e[8];
            if ((rc = io_read(mv, p, e, 8u)) != TC_OK) { goto fail; }
            const uint32_t cnt = ld32(e), d = ld32(e + 4);
            for (uint32_t k = 0; k < cnt && si < mv->n; ++k) {
                mv->pts[si] = acc;
                mv->durs[si] = d;
                acc += d;
                si++;
            }
        }
    }
    if (t.stsz_size != 0u) {
        for (uint32_t i = 0; i < mv->n; ++i) { mv->sizes[i] = t.stsz_size; }
    } else {
        for (uint32_t i = 0; i < mv->n; ++i) {
            uint8_t e[4];
            if ((rc = io_read(mv, t.stsz + (uint64_t)i * 4u, e, 4u)) != TC_OK) { goto fail; }
            mv->sizes[i] = ld32(e);
        }
        for (uint32_t i = 0; i < mv->n; ++i) {
            if (mv->sizes[i] == 0u) { rc = tc_fail(TC_ERR_MALFORMED, "stsz size 0"); goto fail; }
        }
    }
    if ((rc = expand_offsets(mv, &t)) != TC_OK) { goto fail; }
    for (uint32_t i = 0; i < mv->n; ++i) {
        int in = 0;
        for (int k = 0; k < nm && !in; ++k) {
            in = mv->offsets[i] >= mlo[k] && mv->offsets[i] < mhi[k] && mv->sizes[i] <= mhi[k] - mv->offsets[i];
        }
        if (!in) { rc = tc_fail(TC_ERR_MALFORMED, "video sample %u outside every mdat", i); goto fail; }
    }
    if (t.has_stss) {
        for (uint32_t i = 0; i < t.stss_n; ++i) {
            uint8_t e[4];
            if ((rc = io_read(mv, t.stss + (uint64_t)i * 4u, e, 4u)) != TC_OK) { goto fail; }
            const uint32_t s = ld32(e);
            if (s < 1u || s > mv->n) { rc = tc_fail(TC_ERR_MALFORMED, "stss out of range"); goto fail; }
            mv->sync[(s - 1u) / 8u] |= (uint8_t)(1u << ((s - 1u) % 8u));
        }
    } else {
        memset(mv->sync, 0xFF, (size_t)mv->n / 8u + 1u);
    }
    {
        const uint8_t* p = t.tpcc;
        topos_movie_config* c = &mv->cfg;
        c->struct_size = (uint32_t)sizeof(*c);
        c->abi_version = TOPOS_CODEC_ABI_VERSION;
        c->profile = p[6]; c->pixel_format = p[7]; c->bit_depth = p[8]; c->qmatrix_id = p[9]; c->qp_base = p[10];
        c->alpha_mode = p[11]; c->alpha_bit_depth = p[12];
        c->alpha_premultiplied = (p[13] & 1u) != 0u;
        c->color_range = (p[13] & 2u) != 0u;
        c->color_primaries = p[14]; c->color_transfer = p[15]; c->color_matrix = p[16]; c->chroma_siting = p[17];
        c->sar_num = (uint16_t)ld16(p + 18);
        c->sar_den = (uint16_t)ld16(p + 20);
        c->visible_width = (uint16_t)t.entry_w;
        c->visible_height = (uint16_t)t.entry_h;
        topos_frame_config fc;
        movie_to_frame(c, &fc);
        if ((rc = tc_frame_config_validate(&fc)) != TC_OK) { goto fail; }
        mv->width = t.entry_w;
        mv->height = t.entry_h;
        c->timescale = mv->timescale;
        if (c->sar_num == 0u && c->sar_den == 0u) { c->sar_num = 1u; c->sar_den = 1u; }
    }
    if (t.has_tpcb) {
        if (t.tpcb[6] != mv->cfg.alpha_mode || t.tpcb[7] != mv->cfg.alpha_bit_depth) {
            rc = tc_fail(TC_ERR_MALFORMED, "tpcB alpha mode/depth differ from tpcC");
            goto fail;
        }
        if (ld32(t.tpcb + 16) != mv->n) { rc = tc_fail(TC_ERR_MALFORMED, "tpcB frame_count != samples"); goto fail; }
        topos_alpha_budget_info* bi = &mv->budget;
        bi->struct_size = (uint32_t)sizeof(*bi);
        bi->abi_version = TOPOS_CODEC_ABI_VERSION;
        bi->target_ratio_bp = (uint16_t)ld16(t.tpcb + 8);
        bi->actual_ratio_bp = (uint16_t)ld16(t.tpcb + 10);
        bi->max_abs_error = (uint16_t)ld16(t.tpcb + 12);
        bi->flags = (uint16_t)ld16(t.tpcb + 14);
        bi->frame_count = ld32(t.tpcb + 16);
        mv->budget_set = 1u;
    }
    *out = mv;
    return TC_OK;
fail:
    movie_free(mv);
    return rc;
}

/* pread source of tc_movie_open_fd */
static int32_t fd_read(void* ctx, uint64_t off, void* buf, size_t len)
{
    const int fd = (int)(intptr_t)ctx;
    uint8_t* p = (uint8_t*)buf;
    while (len != 0u) {
        const ssize_t r = pread(fd, p, len, (off_t)off);
        if (r < 0 && errno == EINTR) { continue; }
        if (r <= 0) { return TC_ERR_IO; }
        p += r;
        off += (uint64_t)r;
        len -= (size_t)r;
    }
    return TC_OK;
}

/* a read-only descriptor that the caller keeps open for the movie's life
 * (reference topos_codec.h:883-889) */
int32_t tc_movie_open_fd(int fd, uint64_t length, topos_movie** out)
{
    if (fd < 0) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "fd < 0"); }
    topos_io io;
    memset(&io, 0, sizeof(io));
    io.struct_size = (uint32_t)sizeof(io);
    io.abi_version = TOPOS_CODEC_ABI_VERSION;
    io.ctx = (void*)(intptr_t)fd;
    io.read = fd_read;
    io.length = length;
    int32_t rc = tc_movie_open(&io, out);
    if (rc == TC_OK) { (*out)->fd = fd; }
    return rc;
}

int32_t tc_movie_alpha_budget(const topos_movie* mv, topos_alpha_budget_info* info)
{
    if (mv == NULL || info == NULL || info->struct_size != (uint32_t)sizeof(topos_alpha_budget_info) ||
        info->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "info NULL or struct mismatch");
    }
    if (!mv->budget_set) { return tc_fail(TC_ERR_STATE, "the file carries no alpha budget (no tpcB)"); }
    *info = mv->budget;
    return TC_OK;
}

int32_t tc_movie_read_raw_meta(const topos_movie* mv, topos_raw_meta* out)
{
    if (mv == NULL || out == NULL || out->struct_size != (uint32_t)sizeof(topos_raw_meta) ||
        out->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "raw meta NULL or struct mismatch");
    }
    if (!mv->rmeta_set) { return tc_fail(TC_ERR_STATE, "the file carries no trwm"); }
    *out = mv->rmeta;
    return TC_OK;
}

int32_t tc_movie_info(const topos_movie* mv, topos_movie_info* info)
{
    if (mv == NULL || info == NULL || info->struct_size != (uint32_t)sizeof(topos_movie_info) ||
        info->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "info NULL or struct mismatch");
    }
    const topos_movie_config* c = &mv->cfg;
    memset(info->reserved, 0, sizeof(info->reserved));
    info->visible_width = (uint16_t)mv->width;
    info->visible_height = (uint16_t)mv->height;
    info->profile = c->profile; info->pixel_format = c->pixel_format; info->bit_depth = c->bit_depth;
    info->qmatrix_id = c->qmatrix_id; info->qp_base = c->qp_base;
    info->qp_delta_luma = c->qp_delta_luma; info->qp_delta_chroma = c->qp_delta_chroma;
    info->alpha_mode = c->alpha_mode; info->alpha_bit_depth = c->alpha_bit_depth;
    info->alpha_premultiplied = c->alpha_premultiplied; info->color_range = c->color_range;
    info->color_primaries = c->color_primaries; info->color_transfer = c->color_transfer;
    info->color_matrix = c->color_matrix; info->chroma_siting = c->chroma_siting;
    info->sar_num = c->sar_num; info->sar_den = c->sar_den;
    info->timescale = mv->timescale;
    info->sample_count = mv->n;
    info->faststart = mv->faststart;
    info->index_bytes = mv->index_bytes;
    info->reserved[0] = mv->major;
    info->reserved[1] = mv->meta_set ? mv->meta.tier_id : 0u;
    return TC_OK;
}

/* sample header against the track configuration (mov.c:4419-4460) */
static int32_t sample_check(const topos_movie* mv, uint32_t i, const uint8_t* p, size_t size)
{
    if (p == NULL || size < TC_HDR_SIZE) { return tc_fail(TC_ERR_TRUNCATED, "sample %u shorter than a header", i); }
    tc_view* v = (tc_view*)malloc(sizeof(tc_view));
    if (v == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "packet view"); }
    int32_t rc = tc_parse_structure(p, size, v);
    const tc_fhdr fh = v->fh;
    free(v);
    if (rc != TC_OK) { return rc; }
    if (fh.packet_size != size) { return tc_fail(TC_ERR_MALFORMED, "sample %u frame_packet_size != stsz", i); }
    topos_movie_config c = mv->cfg;
    c.visible_width = (uint16_t)mv->width;
    c.visible_height = (uint16_t)mv->height;
    if (fh.major != mv->major || cfg_mismatch(&fh, &c, 1)) {
        return tc_fail(TC_ERR_MALFORMED, "sample %u header differs from tpcC / the sample entry", i);
    }
    return TC_OK;
}

int32_t tc_movie_packet(const topos_movie* mv, uint32_t sample_index, uint8_t* buf, size_t cap, size_t* need_size)
{
    if (mv == NULL || sample_index >= mv->n) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "sample index out of range"); }
    const size_t sz = mv->sizes[sample_index];
    if (need_size != NULL) { *need_size = sz; }
    const uint64_t off = mv->offsets[sample_index];
    if (off > mv->io.length || sz > mv->io.length - off) {
        return tc_fail(TC_ERR_MALFORMED, "sample %u range past the end", sample_index);
    }
    if (buf == NULL || cap < sz) { return TC_ERR_BUFFER_TOO_SMALL; }
    int32_t rc = io_read(mv, off, buf, sz);
    return rc != TC_OK ? rc : sample_check(mv, sample_index, buf, sz);
}

/* parallel pread of one contiguous file range (descriptor sources only:
 * pread needs no shared file position) */
typedef struct pread_part { int fd; uint64_t off; uint8_t* dst; size_t len; int32_t rc; } pread_part;

static void* pread_worker(void* arg)
{
    pread_part* p = (pread_part*)arg;
    p->rc = fd_read((void*)(intptr_t)p->fd, p->off, p->dst, p->len);
    return NULL;
}

#define MOV_READ_THREADS 8
#define MOV_READ_SPLIT ((size_t)4 << 20)

static int32_t read_span(const topos_movie* mv, uint64_t off, uint8_t* dst, size_t len)
{
    if (mv->fd < 0 || len < 2u * MOV_READ_SPLIT) { return io_read(mv, off, dst, len); }
    if (off > mv->io.length || len > mv->io.length - off) { return tc_fail(TC_ERR_TRUNCATED, "read past end"); }
    size_t parts = len / MOV_READ_SPLIT;
    if (parts > MOV_READ_THREADS) { parts = MOV_READ_THREADS; }
    pread_part pp[MOV_READ_THREADS];
    pthread_t th[MOV_READ_THREADS];
    int started[MOV_READ_THREADS] = { 0 };
    const size_t step = (len / parts + 4095u) & ~(size_t)4095u;
    size_t o = 0;
    for (size_t k = 0; k < parts; ++k) {
        const size_t n = k + 1u == parts ? len - o : step;
        pp[k] = (pread_part){ mv->fd, off + o, dst + o, n, TC_OK };
        o += n;
        /* part 0 runs on this thread; a part whose thread cannot start runs here too */
        started[k] = k > 0u && pthread_create(&th[k], NULL, pread_worker, &pp[k]) == 0;
    }
    int32_t rc = TC_OK;
    for (size_t k = 0; k < parts; ++k) {
        if (!started[k]) { pread_worker(&pp[k]); }
    }
    for (size_t k = 0; k < parts; ++k) {
        if (started[k]) { pthread_join(th[k], NULL); }
        if (pp[k].rc != TC_OK && rc == TC_OK) { rc = tc_fail(TC_ERR_IO, "io read failed"); }
    }
    return rc;
}

/* bounds of samples [start, start + count) and their packed layout */
static int32_t batch_layout(const topos_movie* mv, uint32_t start, uint32_t count, size_t* offsets_out,
                            size_t* sizes_out, uint64_t* total_out)
{
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const uint64_t off = mv->offsets[start + i], sz = mv->sizes[start + i];
        if (off > mv->io.length || sz > mv->io.length - off) {
            return tc_fail(TC_ERR_MALFORMED, "sample %u range past the end", start + i);
        }
        if (offsets_out != NULL) { offsets_out[i] = (size_t)total; }
        if (sizes_out != NULL) { sizes_out[i] = (size_t)sz; }
        total += sz;
    }
    if (total > (uint64_t)SIZE_MAX) { return tc_fail(TC_ERR_MALFORMED, "range bytes overflow"); }
    *total_out = total;
    return TC_OK;
}

/* the packed samples into arena (one read when contiguous in the file), then
 * every header checked against the track */
static int32_t batch_fill(const topos_movie* mv, uint32_t start, uint32_t count, uint8_t* arena, uint64_t total)
{
    int contiguous = 1;
    for (uint32_t i = 1; i < count && contiguous; ++i) {
        contiguous = mv->offsets[start + i] == mv->offsets[start + i - 1u] + mv->sizes[start + i - 1u];
    }
    int32_t rc = TC_OK;
    if (contiguous) {
        rc = total == 0u ? TC_OK : read_span(mv, mv->offsets[start], arena, (size_t)total);
    } else {
        size_t o = 0;
        for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
            rc = io_read(mv, mv->offsets[start + i], arena + o, mv->sizes[start + i]);
            o += mv->sizes[start + i];
        }
    }
    size_t o = 0;
    for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
        rc = sample_check(mv, start + i, arena + o, mv->sizes[start + i]);
        o += mv->sizes[start + i];
    }
    return rc;
}

int32_t tc_movie_packet_batch(const topos_movie* mv, uint32_t start, uint32_t count, uint8_t* arena, size_t cap,
                              size_t* offsets_out, size_t* sizes_out, size_t* need_total)
{
    if (mv == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "movie NULL"); }
    if (count == 0u) { return TC_OK; }
    if (start >= mv->n || count > mv->n - start) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "sample range"); }
    uint64_t total = 0;
    int32_t rc = batch_layout(mv, start, count, offsets_out, sizes_out, &total);
    if (rc != TC_OK) { return rc; }
    if (need_total != NULL) { *need_total = (size_t)total; }
    if (arena == NULL || cap < total) { return TC_ERR_BUFFER_TOO_SMALL; }
    return batch_fill(mv, start, count, arena, total);
}

int32_t tc_movie_packet_pts(const topos_movie* mv, uint32_t sample_index, uint64_t* pts_tick, uint32_t* dur_tick)
{
    if (mv == NULL || sample_index >= mv->n || (pts_tick == NULL && dur_tick == NULL)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet_pts arguments");
    }
    if (pts_tick != NULL) { *pts_tick = mv->pts[sample_index]; }
    if (dur_tick != NULL) { *dur_tick = mv->durs[sample_index]; }
    return TC_OK;
}

int32_t tc_movie_packet_sync(const topos_movie* mv, uint32_t sample_index, uint8_t* is_sync)
{
    if (mv == NULL || is_sync == NULL || sample_index >= mv->n) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "packet_sync arguments");
    }
    *is_sync = (uint8_t)((mv->sync[sample_index / 8u] >> (sample_index % 8u)) & 1u);
    return TC_OK;
}

int32_t tc_movie_prev_sync(const topos_movie* mv, uint32_t sample_index, uint32_t* sync_index)
{
    if (mv == NULL || sync_index == NULL || sample_index >= mv->n) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "prev_sync arguments");
    }
    for (uint32_t k = sample_index;; --k) {
        if ((mv->sync[k / 8u] >> (k % 8u)) & 1u) { *sync_index = k; return TC_OK; }
        if (k == 0u) { return tc_fail(TC_ERR_STATE, "no sync sample at or before %u", sample_index); }
    }
}

void tc_movie_close(topos_movie* mv) { movie_free(mv); }

/* new offsets of the moov-first layout: samples in order right after the
 * moov of moov_len bytes and the 16-byte mdat header */
static int assign_offsets(const topos_movie* mv, uint64_t moov_len, uint64_t* noff)
{
    uint64_t acc = (uint64_t)MOV_FTYP_BYTES + moov_len + MOV_MDAT_HDR;
    int big = 0;
    for (uint32_t i = 0; i < mv->n; ++i) {
        noff[i] = acc;
        big |= acc > 0xFFFFFFFFull;
        acc += mv->sizes[i];
    }
    return big;
}

/* moov-first rewrite (mov.c:4809-5005): the offset values do not change the
 * moov length, so the moov is built once with zero offsets to measure it,
 * upgraded to co64 when an offset needs 64 bits, and built again with the
 * final offsets */
int32_t tc_movie_faststart(const topos_io* src, const topos_io* dst)
{
    if (src == NULL || dst == NULL || dst->write == NULL || dst->struct_size != (uint32_t)sizeof(topos_io) ||
        dst->abi_version != TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "faststart arguments");
    }
    topos_movie* mv = NULL;
    int32_t rc = tc_movie_open(src, &mv);
    if (rc != TC_OK) { return rc; }
    if (mv->faststart) { movie_free(mv); return TC_OK; }
    topos_movie_config cfg = mv->cfg;
    cfg.visible_width = (uint16_t)mv->width;
    cfg.visible_height = (uint16_t)mv->height;
    cfg.timescale = mv->timescale;
    const mov_extra x = { mv->budget_set ? &mv->budget : NULL, mv->meta_set ? &mv->meta : NULL,
                          mv->rmeta_set ? &mv->rmeta : NULL };
    uint64_t total = 0, payload = 0;
    for (uint32_t i = 0; i < mv->n; ++i) { total += mv->durs[i]; payload += mv->sizes[i]; }
    uint64_t* noff = (uint64_t*)calloc(mv->n, sizeof(uint64_t));
    uint8_t* sync = (uint8_t*)calloc(mv->n, 1u);
    abuf b = { NULL, 0, 0, 0 };
    if (noff == NULL || sync == NULL) { rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "faststart"); }
    for (uint32_t i = 0; i < mv->n && rc == TC_OK; ++i) { sync[i] = (mv->sync[i / 8u] >> (i % 8u)) & 1u; }
    int co64 = 0;
    if (rc == TC_OK) { rc = moov_build(&b, &cfg, mv->major, mv->n, noff, mv->sizes, mv->durs, sync, total, 0, &x); }
    if (rc == TC_OK) {
        co64 = assign_offsets(mv, b.n, noff);
        for (int round = 0; round < (co64 ? 2 : 1) && rc == TC_OK; ++round) {
            free(b.d);
            memset(&b, 0, sizeof(b));
            rc = moov_build(&b, &cfg, mv->major, mv->n, noff, mv->sizes, mv->durs, sync, total, co64, &x);
            if (rc == TC_OK && co64) { (void)assign_offsets(mv, b.n, noff); }
        }
    }
    if (rc == TC_OK) {
        uint8_t mh[MOV_MDAT_HDR];
        be32(mh, 1u);
        memcpy(mh + 4, "mdat", 4u);
        be64(mh + 8, payload + MOV_MDAT_HDR);
        int32_t w = dst->write(dst->ctx, MOV_FTYP, MOV_FTYP_BYTES);
        if (w == TC_OK) { w = dst->write(dst->ctx, b.d, b.n); }
        if (w == TC_OK) { w = dst->write(dst->ctx, mh, MOV_MDAT_HDR); }
        uint8_t* cp = (uint8_t*)malloc((size_t)1 << 20);
        if (cp == NULL && mv->n != 0u) { w = TC_ERR_OUT_OF_MEMORY; }
        for (uint32_t i = 0; i < mv->n && w == TC_OK; ++i) {
            uint64_t left = mv->sizes[i], o = mv->offsets[i];
            while (left != 0u && w == TC_OK) {
                const size_t ch = left > ((uint64_t)1 << 20) ? ((size_t)1 << 20) : (size_t)left;
                w = src->read(src->ctx, o, cp, ch);
                if (w == TC_OK) { w = dst->write(dst->ctx, cp, ch); }
                o += ch;
                left -= ch;
            }
        }
        free(cp);
        /* a failing callback's own status is returned, as the reference does */
        if (w != TC_OK) { rc = w; }
    }
    free(b.d);
    free(noff);
    free(sync);
    movie_free(mv);
    return rc;
}

/* page-locked staging for tcx_movie_decode_batch_dev: a few grow-only
 * buffers shared by all movies, one per concurrent call (a call that finds
 * every slot busy stages through ordinary memory) */
#define MOV_PIN_SLOTS 4
static struct { void* p; size_t cap; int busy; } g_pin[MOV_PIN_SLOTS];
static pthread_mutex_t g_pin_mu = PTHREAD_MUTEX_INITIALIZER;

static int pin_take(size_t bytes, uint8_t** out)
{
    int slot = -1;
    pthread_mutex_lock(&g_pin_mu);
    for (int k = 0; k < MOV_PIN_SLOTS; ++k) {
        if (!g_pin[k].busy && (slot < 0 || (g_pin[k].cap >= bytes && g_pin[slot].cap < bytes))) { slot = k; }
    }
    if (slot >= 0) { g_pin[slot].busy = 1; }
    pthread_mutex_unlock(&g_pin_mu);
    if (slot < 0) { return -1; }
    if (g_pin[slot].cap < bytes) {
        tcg_pinned_free(g_pin[slot].p);
        const size_t cap = (bytes + ((size_t)1 << 20) - 1u) & ~(((size_t)1 << 20) - 1u);
        g_pin[slot].p = tcg_pinned_alloc(cap);
        g_pin[slot].cap = g_pin[slot].p != NULL ? cap : 0u;
    }
    if (g_pin[slot].p == NULL) {
        pthread_mutex_lock(&g_pin_mu);
        g_pin[slot].busy = 0;
        pthread_mutex_unlock(&g_pin_mu);
        return -1;
    }
    *out = (uint8_t*)g_pin[slot].p;
    return slot;
}

static void pin_give(int slot)
{
    if (slot < 0) { return; }
    pthread_mutex_lock(&g_pin_mu);
    g_pin[slot].busy = 0;
    pthread_mutex_unlock(&g_pin_mu);
}

/* a run of samples straight into one HBM arena and decoded as one batch (no
 * reference counterpart): the packets of samples [start, start + count) are
 * read into page-locked staging (in parallel for descriptor sources), checked,
 * uploaded with one DMA copy, then decoded into dev_planes (count x 4) */
int32_t tcx_movie_decode_batch_dev(const topos_movie* mv, uint32_t start, uint32_t count, uint64_t dev_arena,
                                   size_t arena_cap, const uint64_t* dev_planes, const size_t* dev_strides,
                                   topos_frame_output* infos, int32_t* rcs)
{
    if (mv == NULL || count == 0u || dev_planes == NULL || infos == NULL || rcs == NULL) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "movie batch arguments");
    }
    if (start >= mv->n || count > mv->n - start) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "sample range"); }
    uint64_t total = 0;
    size_t* offs = (size_t*)calloc(count, sizeof(size_t));
    size_t* sizes = (size_t*)calloc(count, sizeof(size_t));
    const uint8_t** hp = (const uint8_t**)calloc(count, sizeof(uint8_t*));
    uint64_t* dp = (uint64_t*)calloc(count, sizeof(uint64_t));
    uint8_t* host = NULL;
    int slot = -1;
    int32_t rc = offs && sizes && hp && dp ? TC_OK : tc_fail(TC_ERR_OUT_OF_MEMORY, "movie batch");
    if (rc == TC_OK) { rc = batch_layout(mv, start, count, offs, sizes, &total); }
    /* the arena comes from tcx_dev_alloc, whose tail slack covers the
     * decoder's read-ahead past the last packet */
    if (rc == TC_OK && total > arena_cap) {
        rc = tc_fail(TC_ERR_BUFFER_TOO_SMALL, "device arena %zu < %llu bytes", arena_cap, (unsigned long long)total);
    }
    if (rc == TC_OK && (slot = pin_take((size_t)total, &host)) < 0 &&
        (host = (uint8_t*)malloc(total ? (size_t)total : 1u)) == NULL) {
        rc = tc_fail(TC_ERR_OUT_OF_MEMORY, "movie batch staging");
    }
    if (rc == TC_OK) { rc = batch_fill(mv, start, count, host, total); }
    if (rc == TC_OK) { rc = tcx_h2d(dev_arena, host, (size_t)total); }
    if (rc == TC_OK) {
        for (uint32_t i = 0; i < count; ++i) {
            hp[i] = host + offs[i];
            dp[i] = dev_arena + offs[i];
        }
        rc = tcx_decode_batch_dev(count, hp, sizes, dp, dev_planes, dev_strides, infos, rcs);
    }
    if (slot >= 0) {
        pin_give(slot);
    } else {
        free(host);
    }
    free(offs);
    free(sizes);
    free(hp);
    free(dp);
    return rc;
}
