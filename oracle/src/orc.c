/* ORACLE — test infrastructure only. A plain, single-threaded C restatement of
 * the Topos intra frame-coding path (rANS2 "V7-R2" colour slices, Rice alpha
 * slices, frame/slice framing, reconstruction and output deblocking). It is
 * the checker the GPU library is compared against; nothing here is shipped.
 *
 * Each function cites the reference routine it restates (paths relative to
 * the reference tree). Pinning: tests/test_oracle_*.py check this code against
 * the reference's golden vectors (native/tests/conformance golden .bin files), against the
 * compiled reference where it is built (oracle/_ref), and against closed forms.
 */
#include "orc.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ utils */

static void be16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)(v >> 8); p[1] = (uint8_t)v; }
static void be32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}
static uint32_t ld16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }
static uint32_t ld32(const uint8_t* p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint32_t bitlen(uint32_t m) { return m == 0u ? 0u : 32u - (uint32_t)__builtin_clz(m); }

/* zlib CRC32, reflected poly 0xEDB88320, bytewise (native/src/common/crc32.c:7-10,105). */
uint32_t orc_crc32(const uint8_t* p, size_t n)
{
    /* per thread: the restatement is called from concurrent threads */
    static _Thread_local uint32_t tab[256];
    static _Thread_local int ready = 0;
    if (!ready) {
        for (uint32_t i = 0; i < 256u; ++i) {
            uint32_t c = i;
            for (int k = 0; k < 8; ++k) { c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1; }
            tab[i] = c;
        }
        ready = 1;
    }
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) { c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8); }
    return c ^ 0xFFFFFFFFu;
}

/* signed <-> non-negative map (native/src/entropy/rice.c:3-17) */
static uint32_t map_signed(int32_t v)
{
    int64_t s = v;
    return s >= 0 ? (uint32_t)((uint64_t)s << 1) : (uint32_t)(((uint64_t)(-s) << 1) - 1u);
}
static int32_t unmap_signed(uint32_t m)
{
    return (m & 1u) == 0u ? (int32_t)(m / 2u) : -(int32_t)(((uint64_t)m + 1u) / 2u);
}

/* DC predictor (native/src/entropy/block_coding.c:7-18) */
static int32_t dc_predict(int has_left, int32_t left, int has_top, int32_t top)
{
    if (has_left && has_top) {
        int64_t s = (int64_t)left + top;
        return s >= 0 ? (int32_t)((s + 1) >> 1) : -(int32_t)((-s) >> 1);
    }
    if (has_left) { return left; }
    if (has_top) { return top; }
    return 0;
}

/* ------------------------------------------------------- transform / quant */

/* 2-pass integer DCT F = M x' M^T on a level-shifted block
 * (native/src/transform/transform.c:18-67, dispatch.c:402-430). For bd <= 12
 * the level-shifted sample is held in int16 exactly as the reference does;
 * for bd > 12 the result is pre-distorted (transform.c:133-150). */
void orc_fdct(const int32_t x_in[64], int32_t bit_depth, int32_t F[64])
{
    int64_t x[64], A[64];
    for (int i = 0; i < 64; ++i) { x[i] = bit_depth <= 12 ? (int64_t)(int16_t)x_in[i] : x_in[i]; }
    for (int y = 0; y < 8; ++y) {
        for (int v = 0; v < 8; ++v) {
            int64_t acc = 0;
            for (int k = 0; k < 8; ++k) { acc += x[y * 8 + k] * orc_M[v * 8 + k]; }
            A[y * 8 + v] = acc;
        }
    }
    for (int u = 0; u < 8; ++u) {
        for (int v = 0; v < 8; ++v) {
            int64_t acc = 0;
            for (int y = 0; y < 8; ++y) { acc += (int64_t)orc_M[u * 8 + y] * A[y * 8 + v]; }
            F[u * 8 + v] = (int32_t)acc;
        }
    }
    if (bit_depth > 12) {
        for (int i = 0; i < 64; ++i) {
            const int64_t w = (int64_t)orc_W[i] * orc_E[i >> 3] * orc_E[i & 7];
            const int64_t num = (int64_t)F[i] * ((int64_t)1 << 32);
            F[i] = num >= 0 ? (int32_t)((2 * num + w) / (2 * w)) : -(int32_t)((-2 * num + w) / (2 * w));
        }
    }
}

/* Q = max(1, (qm * qp_scale[qp] + 128) >> 8) (native/src/transform/quant.c:217-225) */
uint32_t orc_qstep(uint32_t tbl, uint16_t qm, uint32_t qp)
{
    if (qp > 95u) { return 0u; }
    uint32_t q = (uint32_t)(((uint64_t)qm * orc_qp_scale[tbl & 1u][qp] + 128u) >> 8);
    return q == 0u ? 1u : q;
}

/* coefficient clamp: 2^25 for bd <= 12, scaled with the sample range above
 * (native/src/transform/quant.c:269-281, codec/color_store.h:51-57). */
static int64_t f_clamp(uint32_t bd)
{
    if (bd <= 12u) { return (int64_t)1 << 25; }
    return (((int64_t)1 << 25) * (int64_t)((1u << (bd - 1u)) - 1u)) / 2047;
}

/* Deadzone quantiser, natural order (native/src/transform/quant.c:310-328). */
void orc_quant(const int32_t F[64], const uint16_t qm[64], uint32_t qp, uint32_t tbl,
               uint32_t bd, int32_t q[64])
{
    const int64_t fc = f_clamp(bd);
    for (int i = 0; i < 64; ++i) {
        const uint32_t Q = orc_qstep(tbl, qm[i], qp);
        const uint32_t dz = i == 0 ? 0u : Q / 4u;
        int64_t v = F[i];
        if (v > fc) { v = fc; } else if (v < -fc) { v = -fc; }
        const uint32_t mag = (uint32_t)(v < 0 ? -v : v);
        const uint32_t num = mag + (Q >> 1);
        const uint32_t ne = num > dz ? num - dz : 0u;
        const uint32_t qm_ = ne / Q;
        q[i] = v < 0 ? -(int32_t)qm_ : (int32_t)qm_;
    }
}

static int32_t round_shift32(int64_t v)
{
    return v >= 0 ? (int32_t)((v + ((int64_t)1 << 31)) >> 32)
                  : -(int32_t)((-v + ((int64_t)1 << 31)) >> 32);
}

/* Dequantise (clamped) + inverse transform x = round((M^T (F' .* W) M) / 2^32)
 * (native/src/transform/quant.c:348-357, transform.c:69-117). */
void orc_dequant_idct(const int32_t q[64], const uint16_t qm[64], uint32_t qp, uint32_t tbl,
                      uint32_t bd, int32_t xh[64])
{
    const int64_t fc = f_clamp(bd);
    int64_t G[64], B[64];
    for (int i = 0; i < 64; ++i) {
        int64_t v = (int64_t)q[i] * orc_qstep(tbl, qm[i], qp);
        if (v > fc) { v = fc; } else if (v < -fc) { v = -fc; }
        if (v > (1 << 30)) { v = 1 << 30; } else if (v < -(1 << 30)) { v = -(1 << 30); }
        G[i] = v * orc_W[i];
    }
    for (int u = 0; u < 8; ++u) {
        for (int x = 0; x < 8; ++x) {
            int64_t acc = 0;
            for (int v = 0; v < 8; ++v) { acc += G[u * 8 + v] * orc_M[v * 8 + x]; }
            B[u * 8 + x] = acc;
        }
    }
    for (int y = 0; y < 8; ++y) {
        for (int x = 0; x < 8; ++x) {
            int64_t acc = 0;
            for (int u = 0; u < 8; ++u) { acc += (int64_t)orc_M[u * 8 + y] * B[u * 8 + x]; }
            xh[y * 8 + x] = round_shift32(acc);
        }
    }
}

/* ------------------------------------------------------------- rANS model */

typedef struct { uint32_t syms; uint16_t freq[64]; uint16_t cum[65]; uint8_t lut[4096]; } rmodel;

/* count -> 4096-normalised model: floor scaling with a floor of 1 for used
 * symbols, then -1 from the largest freq (lowest index on ties) while over,
 * +1 to the largest remainder while under (native/src/entropy/rans.c:10-84). */
static int model_build(rmodel* m, const uint32_t* count, uint32_t syms)
{
    uint64_t total = 0;
    for (uint32_t s = 0; s < syms; ++s) { total += count[s]; }
    if (total == 0) { return -1; }
    uint32_t freq[64], rem[64], sum = 0;
    for (uint32_t s = 0; s < syms; ++s) {
        const uint64_t sc = (uint64_t)count[s] * 4096u;
        freq[s] = count[s] == 0 ? 0u : (uint32_t)(sc / total);
        rem[s] = (uint32_t)(sc % total);
        if (count[s] != 0 && freq[s] == 0) { freq[s] = 1; }
        sum += freq[s];
    }
    while (sum > 4096u) {
        uint32_t big = UINT32_MAX;
        for (uint32_t s = 0; s < syms; ++s) {
            if (freq[s] > 1u && (big == UINT32_MAX || freq[s] > freq[big])) { big = s; }
        }
        if (big == UINT32_MAX) { break; }
        freq[big]--; rem[big] = 0; sum--;
    }
    while (sum < 4096u) {
        uint32_t best = UINT32_MAX;
        for (uint32_t s = 0; s < syms; ++s) {
            if (count[s] == 0) { continue; }
            if (best == UINT32_MAX || rem[s] > rem[best]) { best = s; }
        }
        if (best == UINT32_MAX) { break; }
        freq[best]++; rem[best] = 0; sum++;
    }
    if (sum != 4096u) { return -1; }
    m->syms = syms;
    uint32_t acc = 0;
    for (uint32_t s = 0; s < syms; ++s) {
        m->freq[s] = (uint16_t)freq[s];
        m->cum[s] = (uint16_t)acc;
        for (uint32_t k = acc; k < acc + freq[s]; ++k) { m->lut[k] = (uint8_t)s; }
        acc += freq[s];
    }
    m->cum[syms] = (uint16_t)acc;
    return 0;
}

/* 8-bit proportional row bytes, round-half-up, used symbols >= 1, empty row
 * -> {1,0,...} (native/src/entropy/rans.c:283-306). */
static void row_bytes(uint8_t* out, uint32_t nsym, const uint32_t* count)
{
    uint64_t total = 0;
    for (uint32_t s = 0; s < nsym; ++s) { total += count[s]; }
    if (total == 0) { memset(out, 0, nsym); out[0] = 1; return; }
    for (uint32_t s = 0; s < nsym; ++s) {
        if (count[s] == 0) { out[s] = 0; continue; }
        uint64_t v = ((uint64_t)count[s] * 255u + total / 2u) / total;
        if (v == 0) { v = 1; }
        if (v > 255) { v = 255; }
        out[s] = (uint8_t)v;
    }
}

static int row_model(const uint8_t* in, uint32_t nsym, rmodel* m)
{
    uint32_t c[64];
    for (uint32_t s = 0; s < nsym; ++s) { c[s] = in[s]; }
    return model_build(m, c, nsym);
}

/* log2 by exponent split + atanh series to z^13 (native/src/codec/ctx_ceiling.c:75-89).
 * The encoder's only floating point; it must round exactly like the reference. */
double orc_log2(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, 8);
    const int e = (int)((bits >> 52) & 0x7FFu) - 1023;
    bits = (bits & ~(0x7FFull << 52)) | (1023ull << 52);
    double m;
    memcpy(&m, &bits, 8);
    const double z = (m - 1.0) / (m + 1.0);
    const double z2 = z * z;
    const double at = z * (1.0 + z2 * (1.0 / 3.0 + z2 * (1.0 / 5.0
        + z2 * (1.0 / 7.0 + z2 * (1.0 / 9.0 + z2 * (1.0 / 11.0
        + z2 * (1.0 / 13.0)))))));
    return (double)e + 2.885390081777926814 * at;
}

/* Exact cost of one candidate context split (native/src/codec/codec.c:2159-2175). */
static double rows_cost(const uint32_t* rows, uint32_t ctx_n, uint32_t nsym)
{
    double bits = 0.0;
    for (uint32_t c = 0; c < ctx_n; ++c) {
        const uint32_t* cnt = rows + (size_t)c * nsym;
        uint8_t rb[64];
        rmodel m;
        row_bytes(rb, nsym, cnt);
        if (row_model(rb, nsym, &m) != 0) { continue; }
        for (uint32_t s = 0; s < nsym; ++s) {
            if (cnt[s] != 0) { bits += (double)cnt[s] * (12.0 - orc_log2((double)m.freq[s])); }
        }
    }
    return bits;
}

static uint32_t pos_ctx(uint32_t pos) { return pos < 8u ? 0u : pos < 16u ? 1u : pos < 32u ? 2u : 3u; }
static uint32_t prev_ctx(uint32_t cat, int has_prev)
{
    if (!has_prev) { return 4u; }
    return cat == 0u ? 0u : cat == 1u ? 1u : cat <= 3u ? 2u : 3u;
}
static uint32_t dcc_ctx(uint32_t cat, int has_prev)
{
    if (!has_prev) { return 0u; }
    return 1u + (cat == 0u ? 0u : cat == 1u ? 1u : cat <= 3u ? 2u : 3u);
}

/* -------------------------------------------------------- byte sink / rANS */

typedef struct { uint8_t* p; size_t n, cap; int oom; } bytes_t;

static void bpush(bytes_t* b, uint8_t v)
{
    if (b->n == b->cap) {
        size_t nc = b->cap ? b->cap * 2 : 4096;
        uint8_t* np = (uint8_t*)realloc(b->p, nc);
        if (!np) { b->oom = 1; return; }
        b->p = np; b->cap = nc;
    }
    b->p[b->n++] = v;
}

#define RANS_L (1u << 23)

static uint32_t r_put(uint32_t x, bytes_t* o, uint32_t f, uint32_t cum)
{
    const uint32_t xmax = (1u << 19) * f;
    while (x >= xmax) { bpush(o, (uint8_t)x); x >>= 8; }
    return ((x / f) << 12) + (x % f) + cum;
}
static uint32_t r_chunk(uint32_t x, bytes_t* o, uint32_t w, uint32_t c)
{
    const uint32_t thr = 1u << (31u - w);
    while (x >= thr) { bpush(o, (uint8_t)x); x >>= 8; }
    return (x << w) | (c & ((1u << w) - 1u));
}
/* cat-1 raw magnitude bits; widths > 23 split as [hi b-16][lo 16]
 * (native/src/entropy/rans.h:99-161, codec.c:2206-2217). */
static uint32_t r_suffix(uint32_t x, bytes_t* o, uint32_t cat, uint32_t m)
{
    uint32_t b = cat > 1u ? cat - 1u : 0u;
    if (b == 0u || b > 27u) { return x; }
    if (b > 23u) { x = r_chunk(x, o, b - 16u, m >> 16); b = 16u; }
    return r_chunk(x, o, b, m);
}

typedef struct { uint32_t x; const uint8_t* buf; size_t size, pos; } rdec;

static int rd_init(rdec* d, const uint8_t* buf, size_t size)
{
    if (size < 4) { return ORC_ETRUNC; }
    d->x = ld32(buf); d->buf = buf; d->size = size; d->pos = 4;
    return d->x < RANS_L ? ORC_EMALFORMED : ORC_OK;
}
static int rd_renorm(rdec* d)
{
    while (d->x < RANS_L) {
        if (d->pos >= d->size) { return ORC_ETRUNC; }
        d->x = (d->x << 8) | d->buf[d->pos++];
    }
    return ORC_OK;
}
static int rd_sym(rdec* d, const rmodel* m, uint32_t* s_out)
{
    const uint32_t slot = d->x & 4095u;
    const uint32_t s = m->lut[slot];
    d->x = m->freq[s] * (d->x >> 12) + slot - m->cum[s];
    *s_out = s;
    return rd_renorm(d);
}
static int rd_chunk(rdec* d, uint32_t w, uint32_t* c)
{
    *c = d->x & ((1u << w) - 1u);
    d->x >>= w;
    return rd_renorm(d);
}
static int rd_suffix(rdec* d, uint32_t cat, uint32_t* m)
{
    uint32_t v = 0;
    if (cat > 1u) {
        const uint32_t b = cat - 1u;
        int rc;
        if (b > 23u) {
            uint32_t lo, hi;
            if ((rc = rd_chunk(d, 16, &lo)) != 0) { return rc; }
            if ((rc = rd_chunk(d, b - 16u, &hi)) != 0) { return rc; }
            v = (hi << 16) | lo;
        } else if ((rc = rd_chunk(d, b, &v)) != 0) {
            return rc;
        }
    }
    *m = cat == 0u ? 0u : (1u << (cat - 1u)) | v;
    return ORC_OK;
}

/* ------------------------------------------------------------- framing */

static uint32_t pad8(uint32_t v) { return (v + 7u) / 8u * 8u; }

/* config -> header fields (native/src/codec/codec.c:280-397) */
static void cfg_to_hdr(const orc_cfg* c, orc_hdr* h)
{
    memset(h, 0, sizeof(*h));
    const uint32_t em = c->reserved[0] & 0xFFu;
    h->flags = c->alpha_premultiplied ? 1u : 0u;
    if (em == 8u && c->reserved[5] != 0u) {
        h->flags |= (uint16_t)(((c->reserved[5] & 1u) ? 4u : 0u) | ((c->reserved[5] & 2u) ? 8u : 0u));
    }
    h->profile = c->profile ? c->profile : (c->pixel_format == 3u ? 7u : 3u);
    h->pixel_format = c->pixel_format;
    h->bit_depth = c->bit_depth ? c->bit_depth : 10u;
    h->color_matrix = c->color_matrix ? c->color_matrix
                      : (h->pixel_format == 2u || h->pixel_format == 3u ? 0u : 1u);
    h->qp_base = c->qp_base;
    h->alpha_mode = c->alpha_mode;
    h->alpha_bit_depth = c->alpha_bit_depth;
    h->coded_w = (uint16_t)pad8(c->visible_width);
    h->coded_h = (uint16_t)pad8(c->visible_height);
    h->visible_w = c->visible_width;
    h->visible_h = c->visible_height;
    h->plane_count = (uint8_t)(c->pixel_format == 3u ? 4u : 3u + (c->alpha_mode ? 1u : 0u));
    h->qmatrix_id = c->qmatrix_id;
    h->color_range = c->color_range;
    h->color_primaries = c->color_primaries ? c->color_primaries : 1u;
    h->color_transfer = c->color_transfer ? c->color_transfer
                        : (c->pixel_format == 3u ? (c->bit_depth == 16u ? 8u : 20u) : 13u);
    h->chroma_siting = c->chroma_siting;
    h->sar_num = c->sar_num;
    h->sar_den = c->sar_den;
    if (em == 8u) { h->major = 7; h->entropy_mode = 7; }
    else if (em == 1u) { h->major = 2; h->entropy_mode = 1; h->codebook_version = 1; }
    else { h->major = 1; }
    if (h->major >= 2u) {
        h->minor = 0;
    } else {
        h->minor = h->bit_depth == 16u ? 6u : h->pixel_format == 3u ? 5u : h->pixel_format == 2u ? 3u
                 : h->pixel_format == 1u ? 2u : h->bit_depth != 10u ? 1u : 0u;
        if (h->qp_base >= 64u && h->minor < 4u) { h->minor = 4u; }
        if (h->color_transfer == 13u && h->minor < 7u) { h->minor = 7u; }
    }
}

/* per-plane geometry (native/src/bitstream/frame_header.c:541-588) */
static void derive_geometry(orc_hdr* h)
{
    if (h->pixel_format == 3u) {
        const uint32_t vw = (h->visible_w + 1u) / 2u, vh = (h->visible_h + 1u) / 2u;
        for (int p = 0; p < 4; ++p) {
            h->pvw[p] = vw; h->pvh[p] = vh; h->pcw[p] = pad8(vw); h->pch[p] = pad8(vh);
            h->bcols[p] = h->pcw[p] / 8u; h->brows[p] = h->pch[p] / 8u;
        }
        return;
    }
    const uint32_t cvw = h->pixel_format == 0u ? (h->visible_w + 1u) / 2u : h->visible_w;
    const uint32_t ccw = h->pixel_format == 0u ? pad8(cvw) : h->coded_w;
    for (int p = 0; p < 4; ++p) {
        const int ch = p == 1 || p == 2;
        h->pcw[p] = ch ? ccw : h->coded_w;
        h->pch[p] = h->coded_h;
        h->pvw[p] = ch ? cvw : h->visible_w;
        h->pvh[p] = h->visible_h;
        h->bcols[p] = h->pcw[p] / 8u;
        h->brows[p] = h->pch[p] / 8u;
    }
}

/* 53-byte big-endian frame header, CRC over bytes 0..48 (frame_header.c:590-631) */
static void hdr_write(const orc_hdr* h, uint8_t* o)
{
    o[0] = 'T'; o[1] = 'P'; o[2] = 'I'; o[3] = 'C';
    be16(o + 4, 53); o[6] = h->major; o[7] = h->minor; be16(o + 8, h->flags);
    o[10] = h->profile; o[11] = h->pixel_format; o[12] = h->bit_depth; o[13] = h->alpha_mode;
    o[14] = h->alpha_bit_depth; o[15] = h->frame_type; be16(o + 16, h->gop_id); o[18] = h->ref_distance;
    be16(o + 19, h->coded_w); be16(o + 21, h->coded_h); be16(o + 23, h->visible_w); be16(o + 25, h->visible_h);
    o[27] = h->plane_count; o[28] = h->qmatrix_id; o[29] = h->qp_base; be16(o + 30, h->slice_count);
    o[32] = h->color_range; o[33] = h->color_primaries; o[34] = h->color_transfer;
    o[35] = h->color_matrix; o[36] = h->chroma_siting; be16(o + 37, h->sar_num); be16(o + 39, h->sar_den);
    be32(o + 41, h->packet_size); o[45] = h->entropy_mode; o[46] = h->codebook_version;
    o[47] = h->coding_mode; o[48] = 0; be32(o + 49, orc_crc32(o, 49));
}

static int hdr_read(const uint8_t* d, size_t size, orc_hdr* h)
{
    if (size < 53u) { return ORC_ETRUNC; }
    if (memcmp(d, "TPIC", 4) != 0) { return ORC_EMALFORMED; }
    if (ld32(d + 49) != orc_crc32(d, 49)) { return ORC_ECRC; }
    memset(h, 0, sizeof(*h));
    h->major = d[6]; h->minor = d[7]; h->flags = (uint16_t)ld16(d + 8);
    h->profile = d[10]; h->pixel_format = d[11]; h->bit_depth = d[12]; h->alpha_mode = d[13];
    h->alpha_bit_depth = d[14]; h->frame_type = d[15]; h->gop_id = (uint16_t)ld16(d + 16);
    h->ref_distance = d[18]; h->coded_w = (uint16_t)ld16(d + 19); h->coded_h = (uint16_t)ld16(d + 21);
    h->visible_w = (uint16_t)ld16(d + 23); h->visible_h = (uint16_t)ld16(d + 25);
    h->plane_count = d[27]; h->qmatrix_id = d[28]; h->qp_base = d[29]; h->slice_count = (uint16_t)ld16(d + 30);
    h->color_range = d[32]; h->color_primaries = d[33]; h->color_transfer = d[34];
    h->color_matrix = d[35]; h->chroma_siting = d[36]; h->sar_num = (uint16_t)ld16(d + 37);
    h->sar_den = (uint16_t)ld16(d + 39); h->packet_size = ld32(d + 41);
    h->entropy_mode = d[45]; h->codebook_version = d[46]; h->coding_mode = d[47];
    if (h->plane_count < 3 || h->plane_count > 4 || h->qmatrix_id > 4 || h->bit_depth < 8 ||
        h->bit_depth > 16 || h->coded_w == 0 || h->coded_h == 0) {
        return ORC_EMALFORMED;
    }
    derive_geometry(h);
    return ORC_OK;
}

static int is_chroma(const orc_hdr* h, uint32_t p) { return h->pixel_format != 3u && (p == 1u || p == 2u); }
static int is_alpha(const orc_hdr* h, uint32_t p) { return p == 3u && h->alpha_mode != 0u; }

/* qp_eff = qp_base + plane delta + 2*(bd-10) for bd >= 12, clamped to 63 (or 95
 * when qp_base >= 64) (native/src/codec/codec.c:230,5437-5453) */
static int32_t qp_eff_of(const orc_hdr* h, const orc_cfg* c, uint32_t p)
{
    int32_t qe = (int32_t)h->qp_base + (is_chroma(h, p) ? c->qp_delta_chroma : c->qp_delta_luma)
               + (h->bit_depth >= 12u ? 2 * ((int32_t)h->bit_depth - 10) : 0);
    if (qe < 0) { qe = 0; }
    const int32_t ceil = h->qp_base >= 64u ? 95 : 63;
    return qe > ceil ? ceil : qe;
}

/* ------------------------------------------------ MSB-first bit writer/reader */

typedef struct { bytes_t b; uint64_t acc; uint32_t nbits; } bitw;

static void bw_put(bitw* w, uint32_t n, uint32_t v)
{
    for (int i = (int)n - 1; i >= 0; --i) {
        w->acc = (w->acc << 1) | ((v >> i) & 1u);
        if (++w->nbits == 8u) { bpush(&w->b, (uint8_t)w->acc); w->acc = 0; w->nbits = 0; }
    }
}
static void bw_flush(bitw* w)
{
    if (w->nbits) { bpush(&w->b, (uint8_t)(w->acc << (8u - w->nbits))); w->acc = 0; w->nbits = 0; }
}

/* Rice code: q = m>>k ones, a zero, k low bits; q > 30 escapes to 31 ones plus a
 * 32-bit literal (native/src/entropy/rice.h:92-116). */
static void rice_put(bitw* w, uint32_t k, uint32_t m)
{
    const uint32_t q = m >> k;
    if (q <= 30u) {
        for (uint32_t i = 0; i < q; ++i) { bw_put(w, 1, 1); }
        bw_put(w, 1, 0);
        if (k) { bw_put(w, k, m & ((1u << k) - 1u)); }
    } else {
        bw_put(w, 31, 0x7FFFFFFFu);
        bw_put(w, 32, m);
    }
}

typedef struct { const uint8_t* p; size_t size; uint64_t pos; } bitr;

static int br_bit(bitr* r, uint32_t* b)
{
    if (r->pos >= (uint64_t)r->size * 8u) { return ORC_ETRUNC; }
    *b = (r->p[r->pos >> 3] >> (7u - (r->pos & 7u))) & 1u;
    r->pos++;
    return ORC_OK;
}
static int br_bits(bitr* r, uint32_t n, uint32_t* v)
{
    if (r->pos + n > (uint64_t)r->size * 8u) { return ORC_ETRUNC; }
    uint32_t x = 0, b = 0;
    for (uint32_t i = 0; i < n; ++i) { br_bit(r, &b); x = (x << 1) | b; }
    *v = x;
    return ORC_OK;
}
/* native/src/entropy/rice.h:160-246 */
static int rice_get(bitr* r, uint32_t k, uint32_t mmax, uint32_t* m)
{
    uint32_t q = 0, b;
    for (;;) {
        if (br_bit(r, &b) != 0) { return ORC_ETRUNC; }
        if (b == 0u) { break; }
        if (++q == 31u) { break; }
    }
    uint64_t v;
    if (q == 31u) {
        uint32_t lit;
        if (br_bits(r, 32, &lit) != 0) { return ORC_ETRUNC; }
        v = lit;
    } else {
        uint32_t rem = 0;
        if (k && br_bits(r, k, &rem) != 0) { return ORC_ETRUNC; }
        v = ((uint64_t)q << k) | rem;
    }
    if (v > mmax) { return ORC_EMALFORMED; }
    *m = (uint32_t)v;
    return ORC_OK;
}

/* floor(log2(mean)) capped at 14 (native/src/codec/codec.c:734-741) */
static uint32_t k_estimate(uint64_t sum, uint64_t cnt)
{
    if (cnt == 0 || sum == 0) { return 0; }
    const uint64_t mean = sum / cnt;
    uint32_t k = 0;
    while (k < 14u && (mean >> (k + 1u)) != 0u) { ++k; }
    return k;
}

/* -------------------------------------------------------- encode: colour */

typedef struct {
    uint32_t* dc_m; uint8_t* npair; uint8_t* run; uint32_t* lvl_m;
    size_t blocks, pairs;
    uint32_t dc_hist[29], run_hist[64], lvl_hist[28];
} toks;

/* One band: FDCT, quant, DC prediction, (run, level) tokens and histograms
 * (native/src/codec/codec.c:1702-1848). Returns ORC_EINVAL on the token
 * domain violation the reference rejects. */
static int band_tokens(const orc_hdr* h, const uint16_t* coded, uint32_t p, uint32_t qp,
                       uint32_t y0, uint32_t bh, toks* t)
{
    const uint32_t cols = h->bcols[p], stride = h->pcw[p];
    const uint16_t* qm = orc_qm[h->qmatrix_id][is_chroma(h, p) ? 1 : 0];
    const uint32_t tbl = (h->flags & 8u) ? 1u : 0u;
    const int32_t mid = 1 << (h->bit_depth - 1);
    /* the encoder saturates against the 16-bit domain limits at every depth */
    const uint32_t dc_max = 0x0FFFFFFFu, lv_max = 0x07FFFFFFu;
    int32_t* rowa = (int32_t*)calloc(cols, 4);
    int32_t* rowb = (int32_t*)calloc(cols, 4);
    int bad = 0;
    t->blocks = (size_t)cols * bh;
    t->dc_m = (uint32_t*)malloc(t->blocks * 4);
    t->npair = (uint8_t*)malloc(t->blocks);
    t->run = (uint8_t*)malloc(t->blocks * 63);
    t->lvl_m = (uint32_t*)malloc(t->blocks * 63 * 4);
    t->pairs = 0;
    memset(t->dc_hist, 0, sizeof(t->dc_hist));
    memset(t->run_hist, 0, sizeof(t->run_hist));
    memset(t->lvl_hist, 0, sizeof(t->lvl_hist));
    size_t bi = 0;
    for (uint32_t by = 0; by < bh; ++by) {
        for (uint32_t bx = 0; bx < cols; ++bx, ++bi) {
            int32_t x[64], F[64], q[64];
            const uint16_t* src = coded + (size_t)(y0 + by) * 8u * stride + bx * 8u;
            for (int r = 0; r < 8; ++r) {
                for (int c = 0; c < 8; ++c) { x[r * 8 + c] = (int32_t)src[(size_t)r * stride + c] - mid; }
            }
            orc_fdct(x, h->bit_depth, F);
            orc_quant(F, qm, qp, tbl, h->bit_depth, q);
            const int32_t pred = dc_predict(bx > 0, bx > 0 ? rowb[bx - 1] : 0, by > 0, rowa[bx]);
            uint32_t dcm = map_signed(q[0] - pred);
            if (dcm > dc_max) { dcm = dc_max; bad = 1; }
            t->dc_m[bi] = dcm;
            t->dc_hist[bitlen(dcm)]++;
            rowb[bx] = q[0];
            uint32_t prev = 0, np = 0;
            for (uint32_t pos = 1; pos < 64; ++pos) {
                const int32_t v = q[orc_zigzag[pos]];
                if (v == 0) { continue; }
                uint32_t lm = map_signed(v);
                if (lm > lv_max) { lm = lv_max; bad = 1; }
                const uint32_t run = pos - prev - 1u;
                t->run[t->pairs] = (uint8_t)run;
                t->lvl_m[t->pairs] = lm;
                t->pairs++;
                t->run_hist[run]++;
                t->lvl_hist[bitlen(lm)]++;
                prev = pos;
                np++;
            }
            t->npair[bi] = (uint8_t)np;
            t->run_hist[63]++;
        }
        int32_t* tmp = rowa; rowa = rowb; rowb = tmp;
    }
    free(rowa);
    free(rowb);
    return bad ? ORC_EINVAL : ORC_OK;
}

static void toks_free(toks* t) { free(t->dc_m); free(t->npair); free(t->run); free(t->lvl_m); }

/* Per-slice rANS2 payload: context joints, exact-cost model choice, flags +
 * row tables, backward encode, [state BE][bytes] flush
 * (native/src/codec/codec.c:2122-2151, 2283-2515; rans.c:165-228). */
static int rans2_payload(const toks* t, bytes_t* out)
{
    uint32_t lvl_pos[4 * 28], lvl_prev[5 * 28], dc_prev[5 * 29];
    memset(lvl_pos, 0, sizeof(lvl_pos));
    memset(lvl_prev, 0, sizeof(lvl_prev));
    memset(dc_prev, 0, sizeof(dc_prev));
    {
        size_t pi = 0;
        uint32_t pdc = 0;
        for (size_t b = 0; b < t->blocks; ++b) {
            const uint32_t dcc = bitlen(t->dc_m[b]);
            dc_prev[(b == 0 ? 0u : dcc_ctx(pdc, 1)) * 29u + dcc]++;
            pdc = dcc;
            uint32_t ppos = 0, pcat = 0;
            for (uint32_t k = 0; k < t->npair[b]; ++k, ++pi) {
                const uint32_t cat = bitlen(t->lvl_m[pi]);
                const uint32_t pos = ppos + t->run[pi] + 1u;
                lvl_pos[pos_ctx(pos) * 28u + cat]++;
                lvl_prev[prev_ctx(pcat, k != 0) * 28u + cat]++;
                pcat = cat;
                ppos = pos;
            }
        }
    }
    uint32_t lvl_model = 0, dc_model = 0;
    double best = rows_cost(t->lvl_hist, 1, 28) + 8.0 * 28.0;
    const double pc = rows_cost(lvl_pos, 4, 28) + 8.0 * (double)(4 * 28);
    if (pc < best) { best = pc; lvl_model = 1; }
    const double vc = rows_cost(lvl_prev, 5, 28) + 8.0 * (double)(5 * 28);
    if (vc < best) { lvl_model = 2; }
    const double dn = rows_cost(t->dc_hist, 1, 29) + 8.0 * 29.0;
    const double dp = rows_cost(dc_prev, 5, 29) + 8.0 * (double)(5 * 29);
    if (dp < dn) { dc_model = 1; }
    const uint32_t lvl_n = lvl_model == 1 ? 4u : lvl_model == 2 ? 5u : 1u;
    const uint32_t dc_n = dc_model ? 5u : 1u;

    static _Thread_local rmodel dcm[5], lvm[5], rnm;   /* per thread: too large for the stack frame's budget */
    bpush(out, (uint8_t)(lvl_model | (dc_model << 2)));
    const uint32_t* dc_rows = dc_model ? dc_prev : t->dc_hist;
    for (uint32_t c = 0; c < dc_n; ++c) {
        uint8_t rb[29];
        row_bytes(rb, 29, dc_rows + c * 29u);
        if (row_model(rb, 29, &dcm[c]) != 0) { return ORC_EINVAL; }
        for (int i = 0; i < 29; ++i) { bpush(out, rb[i]); }
    }
    {
        uint8_t rb[64];
        row_bytes(rb, 64, t->run_hist);
        if (row_model(rb, 64, &rnm) != 0) { return ORC_EINVAL; }
        for (int i = 0; i < 64; ++i) { bpush(out, rb[i]); }
    }
    const uint32_t* lv_rows = lvl_model == 1 ? lvl_pos : lvl_model == 2 ? lvl_prev : t->lvl_hist;
    for (uint32_t c = 0; c < lvl_n; ++c) {
        uint8_t rb[28];
        row_bytes(rb, 28, lv_rows + c * 28u);
        if (row_model(rb, 28, &lvm[c]) != 0) { return ORC_EINVAL; }
        for (int i = 0; i < 28; ++i) { bpush(out, rb[i]); }
    }

    bytes_t em = { 0 };
    uint32_t x = RANS_L;
    size_t pi = t->pairs;
    for (size_t b = t->blocks; b-- > 0;) {
        const uint32_t np = t->npair[b];
        pi -= np;
        x = r_put(x, &em, rnm.freq[63], rnm.cum[63]);
        uint8_t posb[63];
        uint32_t pp = 0;
        for (uint32_t k = 0; k < np; ++k) { pp += t->run[pi + k] + 1u; posb[k] = (uint8_t)pp; }
        for (uint32_t k = np; k-- > 0;) {
            const uint32_t lm = t->lvl_m[pi + k];
            const uint32_t ls = bitlen(lm);
            x = r_suffix(x, &em, ls, lm);
            uint32_t lc = 0;
            if (lvl_model == 1) { lc = pos_ctx(posb[k]); }
            else if (lvl_model == 2) { lc = prev_ctx(k ? bitlen(t->lvl_m[pi + k - 1]) : 0u, k != 0); }
            x = r_put(x, &em, lvm[lc].freq[ls], lvm[lc].cum[ls]);
            const uint32_t r = t->run[pi + k];
            x = r_put(x, &em, rnm.freq[r], rnm.cum[r]);
        }
        const uint32_t d = t->dc_m[b];
        const uint32_t ds = bitlen(d);
        x = r_suffix(x, &em, ds, d);
        const uint32_t dctx = dc_model ? dcc_ctx(b ? bitlen(t->dc_m[b - 1]) : 0u, b != 0) : 0u;
        x = r_put(x, &em, dcm[dctx].freq[ds], dcm[dctx].cum[ds]);
    }
    bpush(out, (uint8_t)(x >> 24)); bpush(out, (uint8_t)(x >> 16));
    bpush(out, (uint8_t)(x >> 8)); bpush(out, (uint8_t)x);
    for (size_t i = em.n; i-- > 0;) { bpush(out, em.p[i]); }
    free(em.p);
    return em.oom || out->oom ? ORC_ENOMEM : ORC_OK;
}

/* ------------------------------------------ V1 Rice / V2 VLC colour slices */

/* V2 frozen code lengths per book (codebook_version 1): DC_CAT 29 symbols,
 * RUN 64, LEVEL_CAT 28 with category 0 outside the alphabet
 * (native/src/entropy/vlc_tables.h:28-67). */
static const uint8_t vlc_dc_len[4][29] = {
    { 3, 13, 13, 13, 13, 13, 13, 12, 8, 6, 6, 5, 4, 3, 3, 3, 4, 4, 4, 4, 5, 5, 6, 6, 7, 7, 7, 9, 10 },
    { 3, 10, 15, 15, 14, 14, 11, 7, 6, 6, 5, 4, 3, 3, 3, 3, 4, 4, 5, 5, 6, 7, 7, 7, 7, 8, 9, 12, 14 },
    { 3, 10, 13, 11, 8, 7, 6, 5, 4, 3, 3, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 7, 7, 9, 12, 15, 15, 15, 15 },
    { 3, 6, 5, 4, 4, 3, 3, 3, 4, 4, 4, 5, 5, 5, 6, 7, 7, 7, 8, 9, 10, 11, 14, 14, 14, 14, 14, 14, 13 },
};
static const uint8_t vlc_run_len[4][64] = {
    { 1, 2, 5, 6, 7, 6, 7, 7, 8, 7, 7, 8, 9, 7, 8, 13, 15, 10, 16, 20, 20, 12, 12, 14, 13, 20, 17, 20, 15, 20, 20, 20,
      14, 20, 20, 20, 20, 20, 18, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 3 },
    { 1, 2, 5, 6, 7, 6, 7, 8, 7, 7, 7, 8, 9, 7, 8, 13, 15, 10, 16, 20, 20, 12, 12, 13, 13, 20, 17, 20, 16, 20, 20, 20,
      16, 20, 20, 20, 20, 20, 18, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 3 },
    { 1, 2, 4, 5, 7, 7, 8, 9, 10, 9, 9, 10, 11, 9, 10, 15, 16, 12, 18, 19, 19, 14, 14, 16, 16, 20, 20, 20, 19, 20, 20,
      20, 19, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 20, 3 },
    { 1, 2, 4, 5, 5, 6, 6, 7, 8, 8, 8, 9, 9, 9, 9, 11, 11, 10, 11, 12, 12, 12, 13, 13, 13, 13, 14, 15, 15, 16, 16, 15,
      16, 17, 17, 18, 19, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20, 20,
      20, 20, 20, 4 },
};
static const uint8_t vlc_lvl_len[4][28] = {
    { 0, 10, 9, 8, 7, 5, 4, 4, 3, 3, 3, 3, 3, 4, 5, 5, 6, 6, 7, 8, 8, 9, 8, 7, 7, 7, 9, 10 },
    { 0, 6, 5, 4, 4, 3, 3, 3, 3, 4, 4, 4, 5, 5, 6, 6, 7, 8, 8, 8, 7, 7, 7, 9, 10, 11, 12, 12 },
    { 0, 4, 3, 3, 3, 3, 3, 4, 4, 5, 5, 5, 6, 6, 7, 8, 7, 7, 6, 7, 7, 9, 10, 13, 13, 12, 12, 12 },
    { 0, 3, 2, 3, 3, 3, 4, 5, 5, 6, 6, 6, 6, 6, 6, 6, 7, 8, 9, 10, 13, 13, 13, 13, 13, 13, 13, 13 },
};

const uint8_t* orc_vlc_lengths(uint32_t family, uint32_t book)
{
    if (book > 3u) { return NULL; }
    return family == 0u ? vlc_dc_len[book] : family == 1u ? vlc_run_len[book]
         : family == 2u ? vlc_lvl_len[book] : NULL;
}

/* canonical code: lengths ascending, symbols ascending within one length,
 * consecutive code values (native/src/entropy/vlc.c:37-53) */
typedef struct {
    uint32_t nsym, code[64], first[21], count[21], idx[21];
    uint8_t len[64], order[64];
} vbook;

static void vbook_make(vbook* b, const uint8_t* len, uint32_t nsym)
{
    memset(b, 0, sizeof(*b));
    b->nsym = nsym;
    uint32_t code = 0, n = 0;
    for (uint32_t l = 1; l <= 20u; ++l) {
        b->first[l] = code;
        b->idx[l] = n;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (len[s] != l) { continue; }
            b->len[s] = (uint8_t)l;
            b->code[s] = code++;
            b->order[n++] = (uint8_t)s;
            b->count[l]++;
        }
        code <<= 1;
    }
}

uint32_t orc_vlc_code(uint32_t family, uint32_t book, uint32_t sym)
{
    const uint8_t* len = orc_vlc_lengths(family, book);
    vbook b;
    vbook_make(&b, len, family == 0u ? 29u : family == 1u ? 64u : 28u);
    return b.code[sym];
}

/* one symbol, bit by bit (native/src/entropy/vlc.c:108-126) */
static int vbook_get(bitr* r, const vbook* b, uint32_t* sym)
{
    uint32_t code = 0, bit;
    for (uint32_t l = 1; l <= 20u; ++l) {
        if (br_bit(r, &bit) != 0) { return ORC_ETRUNC; }
        code = (code << 1) | bit;
        if (b->count[l] && code >= b->first[l] && code - b->first[l] < b->count[l]) {
            *sym = b->order[b->idx[l] + code - b->first[l]];
            return ORC_OK;
        }
    }
    return ORC_EMALFORMED;
}

/* category symbol + (cat - 1) suffix bits, m = 1 << (cat-1) | suffix
 * (native/src/entropy/vlc.c:130-151) */
static int vbook_get_cat(bitr* r, const vbook* b, uint32_t mmax, uint32_t* m)
{
    uint32_t cat;
    int rc = vbook_get(r, b, &cat);
    if (rc) { return rc; }
    uint32_t v = 0;
    if (cat > 1u) {
        uint32_t sfx;
        if (br_bits(r, cat - 1u, &sfx) != 0) { return ORC_ETRUNC; }
        v = (1u << (cat - 1u)) | sfx;
    } else if (cat == 1u) {
        v = 1u;
    }
    if (v > mmax) { return ORC_EMALFORMED; }
    *m = v;
    return ORC_OK;
}

static void vbook_put_cat(bitw* w, const vbook* b, uint32_t m)
{
    const uint32_t cat = bitlen(m);
    bw_put(w, b->len[cat], b->code[cat]);
    if (cat > 1u) { bw_put(w, cat - 1u, m & ((1u << (cat - 1u)) - 1u)); }
}

/* book with the fewest total bits (code length + category suffix), ties to
 * the lowest id (native/src/codec/codec.c:1896-1922) */
static uint32_t vlc_pick(const uint8_t* base, uint32_t stride, uint32_t nsym, const uint32_t* hist, int is_cat)
{
    uint64_t best = UINT64_MAX;
    uint32_t id = 0;
    for (uint32_t bk = 0; bk < 4u; ++bk) {
        const uint8_t* len = base + (size_t)bk * stride;
        uint64_t total = 0;
        for (uint32_t s = 0; s < nsym; ++s) {
            if (hist[s]) { total += (uint64_t)hist[s] * (len[s] + (is_cat && s > 1u ? s - 1u : 0u)); }
        }
        if (total < best) { best = total; id = bk; }
    }
    return id;
}

/* V1 colour payload: Rice k per family from the band means, then per block
 * DC (k1), (run k3, level k2) pairs and the end-of-block run 63
 * (native/src/codec/codec.c:1837-1839,1854-1881) */
static int v1_payload(const toks* t, bytes_t* out, uint8_t* k1, uint8_t* k2, uint8_t* k3)
{
    uint64_t sdc = 0, slv = 0, srun = 0;
    for (size_t b = 0; b < t->blocks; ++b) { sdc += t->dc_m[b]; }
    for (size_t i = 0; i < t->pairs; ++i) { slv += t->lvl_m[i]; srun += t->run[i]; }
    srun += 63u * (uint64_t)t->blocks;
    const uint32_t a = k_estimate(sdc, t->blocks), l = k_estimate(slv, t->pairs);
    const uint32_t r = k_estimate(srun, t->pairs + t->blocks);
    bitw w = { { 0 }, 0, 0 };
    size_t pi = 0;
    for (size_t b = 0; b < t->blocks; ++b) {
        rice_put(&w, a, t->dc_m[b]);
        for (uint32_t k = 0; k < t->npair[b]; ++k, ++pi) {
            rice_put(&w, r, t->run[pi]);
            rice_put(&w, l, t->lvl_m[pi]);
        }
        rice_put(&w, r, 63u);
    }
    bw_flush(&w);
    *out = w.b;
    *k1 = (uint8_t)a; *k2 = (uint8_t)l; *k3 = (uint8_t)r;
    return out->oom ? ORC_ENOMEM : ORC_OK;
}

/* V2 colour payload: same symbols as V1 with the frozen canonical books,
 * one book per family chosen by exact bit count
 * (native/src/codec/codec.c:1950-2015) */
static int v2_payload(const toks* t, bytes_t* out, uint8_t* k1, uint8_t* k2, uint8_t* k3)
{
    const uint32_t dbk = vlc_pick(&vlc_dc_len[0][0], 29, 29, t->dc_hist, 1);
    const uint32_t lbk = vlc_pick(&vlc_lvl_len[0][0], 28, 28, t->lvl_hist, 1);
    const uint32_t rbk = vlc_pick(&vlc_run_len[0][0], 64, 64, t->run_hist, 0);
    vbook db, lb, rb;
    vbook_make(&db, vlc_dc_len[dbk], 29);
    vbook_make(&lb, vlc_lvl_len[lbk], 28);
    vbook_make(&rb, vlc_run_len[rbk], 64);
    bitw w = { { 0 }, 0, 0 };
    size_t pi = 0;
    for (size_t b = 0; b < t->blocks; ++b) {
        vbook_put_cat(&w, &db, t->dc_m[b]);
        for (uint32_t k = 0; k < t->npair[b]; ++k, ++pi) {
            bw_put(&w, rb.len[t->run[pi]], rb.code[t->run[pi]]);
            vbook_put_cat(&w, &lb, t->lvl_m[pi]);
        }
        bw_put(&w, rb.len[63], rb.code[63]);
    }
    bw_flush(&w);
    *out = w.b;
    *k1 = (uint8_t)dbk; *k2 = (uint8_t)lbk; *k3 = (uint8_t)rbk;
    return out->oom ? ORC_ENOMEM : ORC_OK;
}

/* -------------------------------------------------------- encode: alpha */

static int32_t med(int32_t a, int32_t b, int32_t c)
{
    if (c >= a && c >= b) { return a < b ? a : b; }
    if (c <= a && c <= b) { return a < b ? b : a; }
    return a + b - c;
}

/* MED residuals inside the band + (run, level) Rice pairs with a trailing
 * (run, 0) terminator (native/src/codec/codec.c:2969-3013,
 * native/src/entropy/block_coding.c:456-481). */
static int alpha_payload(const orc_hdr* h, const uint16_t* coded, uint32_t y0, uint32_t bh,
                         bytes_t* out, uint8_t* k1o, uint8_t* k2o)
{
    const uint32_t w = h->pcw[3], rows = bh * 8u;
    const size_t n = (size_t)w * rows;
    int32_t* res = (int32_t*)malloc(n * 4);
    uint64_t sl = 0, cl = 0, sr = 0, cr = 0, run = 0;
    for (uint32_t y = 0; y < rows; ++y) {
        const uint16_t* row = coded + (size_t)(y0 * 8u + y) * w;
        const uint16_t* up = y ? row - w : row;
        for (uint32_t x = 0; x < w; ++x) {
            int32_t pred;
            if (y == 0) { pred = x == 0 ? 0 : row[x - 1]; }
            else if (x == 0) { pred = up[0]; }
            else { pred = med(row[x - 1], up[x], up[x - 1]); }
            const int32_t r = (int32_t)row[x] - pred;
            res[(size_t)y * w + x] = r;
            if (r != 0) { sr += run; cr++; sl += map_signed(r); cl++; run = 0; }
            else { run++; }
        }
    }
    if (run > 0) { sr += run; cr++; }
    const uint32_t k1 = k_estimate(sl, cl), k2 = k_estimate(sr, cr);
    bitw bw = { { 0 }, 0, 0 };
    size_t pos = 0;
    uint64_t rl = 0;
    for (size_t i = 0; i < n; ++i) {
        if (res[i] == 0) { rl++; continue; }
        rice_put(&bw, k2, (uint32_t)rl);
        rice_put(&bw, k1, map_signed(res[i]));
        pos = i + 1;
        rl = 0;
    }
    if (n - pos > 0) { rice_put(&bw, k2, (uint32_t)(n - pos)); rice_put(&bw, k1, 0); }
    bw_flush(&bw);
    for (size_t i = 0; i < bw.b.n; ++i) { bpush(out, bw.b.p[i]); }
    free(bw.b.p);
    free(res);
    *k1o = (uint8_t)k1;
    *k2o = (uint8_t)k2;
    return ORC_OK;
}

/* -------------------------------------------------------- frame encode */

/* Full frame: pad planes by edge replication (transform/plane.c:19-44), encode
 * every band of every plane, assemble [53 B header][17 B slice header +
 * payload]* in plane, band order (native/src/codec/codec.c:5640-5693,5810-5878). */
int32_t orc_frame_encode(const orc_cfg* cfg, const uint16_t* const planes[4],
                         const size_t strides[4], uint8_t* out, size_t cap, size_t* size_out)
{
    const uint32_t em = cfg->reserved[0] & 0xFFu;
    if (em != 8u && em != 0u && em != 1u) { return ORC_ENOTIMPL; }
    orc_hdr h;
    cfg_to_hdr(cfg, &h);
    derive_geometry(&h);
    const uint32_t sr = cfg->slice_rows ? cfg->slice_rows : 16u;
    uint32_t total = 0, bands[4];
    for (uint32_t p = 0; p < h.plane_count; ++p) {
        bands[p] = (h.brows[p] + sr - 1u) / sr;
        if (bands[p] == 0) { bands[p] = 1; }
        total += bands[p];
    }
    if (total > 512u) { return ORC_ELIMIT; }
    h.slice_count = (uint16_t)total;

    bytes_t pk = { 0 };
    for (int i = 0; i < 53; ++i) { bpush(&pk, 0); }
    int rc = ORC_OK;
    for (uint32_t p = 0; p < h.plane_count && rc == ORC_OK; ++p) {
        const uint32_t cw = h.pcw[p], chh = h.pch[p], vw = h.pvw[p], vh = h.pvh[p];
        const size_t st = strides && strides[p] ? strides[p] : vw;
        uint16_t* coded = (uint16_t*)malloc((size_t)cw * chh * 2);
        for (uint32_t y = 0; y < chh; ++y) {
            const uint16_t* s = planes[p] + (size_t)(y < vh ? y : vh - 1u) * st;
            for (uint32_t x = 0; x < cw; ++x) { coded[(size_t)y * cw + x] = s[x < vw ? x : vw - 1u]; }
        }
        if (is_alpha(&h, p) && h.alpha_mode == 2u) {
            const uint32_t s = 16u - h.alpha_bit_depth, half = 1u << (s - 1u), top = (0xFFFFu >> s) << s;
            for (size_t i = 0; i < (size_t)cw * chh; ++i) {
                uint32_t q = ((coded[i] + half) >> s) << s;
                coded[i] = (uint16_t)(q > top ? top : q);
            }
        }
        const int32_t qe = is_alpha(&h, p) ? 0 : qp_eff_of(&h, cfg, p);
        for (uint32_t y0 = 0; y0 < h.brows[p] && rc == ORC_OK; y0 += sr) {
            const uint32_t bh = h.brows[p] - y0 < sr ? h.brows[p] - y0 : sr;
            bytes_t pay = { 0 };
            orc_slice s = { 0 };
            s.plane = (uint8_t)p; s.y0 = (uint16_t)y0; s.h = (uint16_t)bh;
            if (is_alpha(&h, p)) {
                s.qpd = 64;
                rc = alpha_payload(&h, coded, y0, bh, &pay, &s.k1, &s.k2);
            } else {
                toks t;
                s.qpd = (uint8_t)(64 + qe - (int32_t)h.qp_base);
                rc = band_tokens(&h, coded, p, (uint32_t)qe, y0, bh, &t);
                if (rc == ORC_OK) {
                    rc = em == 8u ? rans2_payload(&t, &pay)
                       : em == 1u ? v2_payload(&t, &pay, &s.k1, &s.k2, &s.k3)
                                  : v1_payload(&t, &pay, &s.k1, &s.k2, &s.k3);
                }
                toks_free(&t);
            }
            if (rc == ORC_OK) {
                uint8_t sh[17];
                be32(sh, (uint32_t)pay.n); sh[4] = s.plane; be16(sh + 5, s.y0); be16(sh + 7, s.h);
                sh[9] = s.qpd; sh[10] = s.k1; sh[11] = s.k2; sh[12] = s.k3;
                be32(sh + 13, orc_crc32(pay.p, pay.n));
                for (int i = 0; i < 17; ++i) { bpush(&pk, sh[i]); }
                for (size_t i = 0; i < pay.n; ++i) { bpush(&pk, pay.p[i]); }
            }
            free(pay.p);
        }
        free(coded);
    }
    if (rc == ORC_OK && pk.oom) { rc = ORC_ENOMEM; }
    if (rc == ORC_OK && pk.n > 268435456u) { rc = ORC_ELIMIT; }
    if (rc == ORC_OK) {
        h.packet_size = (uint32_t)pk.n;
        hdr_write(&h, pk.p);
        if (size_out) { *size_out = pk.n; }
        if (pk.n > cap) { rc = ORC_EBUF; }
        else if (out) { memcpy(out, pk.p, pk.n); }
    }
    free(pk.p);
    return rc;
}

/* -------------------------------------------------------- frame decode */

typedef struct {
    uint16_t* dst; size_t stride; uint32_t vis_w, vis_h;
    uint32_t limit;   /* highest kept scan position (63: all); reduced decode */
} plane_out;

/* dequantise + inverse transform one block and store its visible samples
 * (native/src/codec/color_store.h:454-718) */
static void store_block(const plane_out* po, const int32_t blk[64], const uint16_t* qm, uint32_t qp, uint32_t tbl,
                        uint32_t bd, uint32_t bx, uint32_t brow)
{
    const int32_t mid = 1 << (bd - 1), maxv = (1 << bd) - 1;
    int32_t xh[64];
    if (po->limit < 63u) {
        /* reduced decode keeps scan positions 0..limit (native/src/codec/
         * color_store.h:723-757, native/src/entropy/color_scan.h:395-420) */
        int32_t low[64];
        for (int i = 0; i < 64; ++i) { low[i] = 0; }
        for (uint32_t k = 0; k <= po->limit; ++k) { low[orc_zigzag[k]] = blk[orc_zigzag[k]]; }
        orc_dequant_idct(low, qm, qp, tbl, bd, xh);
    } else {
        orc_dequant_idct(blk, qm, qp, tbl, bd, xh);
    }
    const uint32_t px0 = bx * 8u, py0 = brow * 8u;
    for (uint32_t yy = 0; yy < 8u && py0 + yy < po->vis_h; ++yy) {
        for (uint32_t xx = 0; xx < 8u && px0 + xx < po->vis_w; ++xx) {
            int32_t v = xh[yy * 8 + xx] + mid;
            v = v < 0 ? 0 : v > maxv ? maxv : v;
            po->dst[(size_t)(py0 + yy) * po->stride + px0 + xx] = (uint16_t)v;
        }
    }
}

/* V1 Rice (major 1) or V2 canonical VLC (major 2, entropy mode 1) colour
 * slice: DC residual, (run, level) pairs up to the end-of-block run 63, the
 * 12-bit symbol domains, then byte alignment inside the payload and exact
 * consumption (native/src/bitstream/slice_codec.c:233-341,
 * native/src/entropy/block_coding.c:78-111,146-184). */
static int decode_colour_bits(const orc_hdr* h, const orc_slice* s, const uint8_t* pl, const plane_out* po)
{
    const int vlc = h->major == 2u;
    vbook db, lb, rb;
    if (vlc) {
        if (s->k1 > 3u || s->k2 > 3u || s->k3 > 3u) { return ORC_EMALFORMED; }
        vbook_make(&db, vlc_dc_len[s->k1], 29);
        vbook_make(&lb, vlc_lvl_len[s->k2], 28);
        vbook_make(&rb, vlc_run_len[s->k3], 64);
    } else if (s->k1 > 14u || s->k2 > 14u || s->k3 > 14u) {
        return ORC_EINVAL;
    }
    bitr br = { pl, s->size, 0 };
    const uint32_t p = s->plane, cols = h->bcols[p];
    const uint16_t* qm = orc_qm[h->qmatrix_id][is_chroma(h, p) ? 1 : 0];
    const uint32_t tbl = (h->flags & 8u) ? 1u : 0u;
    const uint32_t qp = (uint32_t)((int32_t)h->qp_base + (int32_t)s->qpd - 64);
    const int64_t dlim = (int64_t)1 << 25;
    int32_t* prow = (int32_t*)calloc(cols, 4);
    int32_t* crow = (int32_t*)calloc(cols, 4);
    int rc = 0;
    for (uint32_t by = 0; by < s->h && rc == 0; ++by) {
        for (uint32_t bx = 0; bx < cols && rc == 0; ++bx) {
            int32_t blk[64];
            memset(blk, 0, sizeof(blk));
            uint32_t m;
            rc = vlc ? vbook_get_cat(&br, &db, 0x08000000u, &m) : rice_get(&br, s->k1, 0x08000000u, &m);
            if (rc) { break; }
            const int32_t pred = dc_predict(bx != 0, bx ? crow[bx - 1] : 0, by != 0, prow[bx]);
            const int64_t dv = (int64_t)pred + unmap_signed(m);
            if (dv > dlim || dv < -dlim) { rc = ORC_EMALFORMED; break; }
            blk[0] = (int32_t)dv;
            uint32_t pos = 1;
            for (;;) {
                uint32_t run;
                rc = vlc ? vbook_get(&br, &rb, &run) : rice_get(&br, s->k3, 63u, &run);
                if (rc || run == 63u) { break; }
                const uint32_t idx = pos + run;
                if (idx > 63u) { rc = ORC_EMALFORMED; break; }
                rc = vlc ? vbook_get_cat(&br, &lb, 0x04000000u, &m) : rice_get(&br, s->k2, 0x04000000u, &m);
                if (rc) { break; }
                blk[orc_zigzag[idx]] = unmap_signed(m);
                pos = idx + 1u;
            }
            if (rc) { break; }
            crow[bx] = blk[0];
            store_block(po, blk, qm, qp, tbl, h->bit_depth, bx, s->y0 + by);
        }
        int32_t* tmp = prow; prow = crow; crow = tmp;
    }
    free(prow);
    free(crow);
    if (rc == 0) {
        const uint64_t aligned = (br.pos + 7u) & ~(uint64_t)7u;
        if (aligned > (uint64_t)s->size * 8u) { rc = ORC_ETRUNC; }
        else if (aligned != (uint64_t)s->size * 8u) { rc = ORC_EMALFORMED; }
    }
    return rc;
}

/* rANS2 colour slice decode + reconstruction into the caller plane
 * (native/src/bitstream/slice_codec.c:554-736; codec/color_store.h:454-718). */
static int decode_colour(const orc_hdr* h, const orc_slice* s, const uint8_t* pl, const plane_out* po)
{
    const size_t n = s->size;
    if (n < 1u + 121u + 4u) { return ORC_ETRUNC; }
    const uint8_t flags = pl[0];
    const uint32_t lvl_model = flags & 3u, dc_model = (flags >> 2) & 1u;
    if ((flags & 0xF8u) != 0u || lvl_model > 2u) { return ORC_EMALFORMED; }
    const uint32_t lvl_n = lvl_model == 1 ? 4u : lvl_model == 2 ? 5u : 1u;
    const uint32_t dc_n = dc_model ? 5u : 1u;
    const size_t prefix = 1u + dc_n * 29u + 64u + lvl_n * 28u;
    if (n < prefix + 4u) { return ORC_ETRUNC; }
    static _Thread_local rmodel dcm[5], lvm[5], rnm;
    size_t off = 1;
    for (uint32_t c = 0; c < dc_n; ++c, off += 29) { if (row_model(pl + off, 29, &dcm[c])) { return ORC_EMALFORMED; } }
    if (row_model(pl + off, 64, &rnm)) { return ORC_EMALFORMED; }
    off += 64;
    for (uint32_t c = 0; c < lvl_n; ++c, off += 28) { if (row_model(pl + off, 28, &lvm[c])) { return ORC_EMALFORMED; } }
    rdec d;
    int rc = rd_init(&d, pl + prefix, n - prefix);
    if (rc) { return rc; }
    const uint32_t p = s->plane, cols = h->bcols[p];
    const uint16_t* qm = orc_qm[h->qmatrix_id][is_chroma(h, p) ? 1 : 0];
    const uint32_t tbl = (h->flags & 8u) ? 1u : 0u;
    const uint32_t qp = (uint32_t)((int32_t)h->qp_base + (int32_t)s->qpd - 64);
    const uint32_t bd = h->bit_depth;
    const uint32_t dc_max = bd > 12u ? 0x0FFFFFFFu : 0x08000000u;
    const uint32_t lv_max = bd > 12u ? 0x07FFFFFFu : 0x04000000u;
    const int64_t dlim = bd > 12u ? ((int64_t)1 << 27) : ((int64_t)1 << 25);
    int32_t* prow = (int32_t*)calloc(cols, 4);
    int32_t* crow = (int32_t*)calloc(cols, 4);
    uint32_t pdc = 0;
    for (uint32_t by = 0; by < s->h && rc == 0; ++by) {
        for (uint32_t bx = 0; bx < cols && rc == 0; ++bx) {
            int32_t blk[64];
            memset(blk, 0, sizeof(blk));
            const uint32_t dctx = (by == 0 && bx == 0) ? 0u : dcc_ctx(pdc, 1);
            uint32_t cat, m;
            if ((rc = rd_sym(&d, &dcm[dc_model ? dctx : 0u], &cat))) { break; }
            if ((rc = rd_suffix(&d, cat, &m))) { break; }
            if (m > dc_max) { rc = ORC_EMALFORMED; break; }
            const int32_t pred = dc_predict(bx != 0, bx ? crow[bx - 1] : 0, by != 0, prow[bx]);
            const int64_t dv = (int64_t)pred + unmap_signed(m);
            if (dv > dlim || dv < -dlim) { rc = ORC_EMALFORMED; break; }
            blk[0] = (int32_t)dv;
            pdc = cat;
            uint32_t pos = 1, pcat = 0;
            int has_prev = 0;
            for (;;) {
                uint32_t rs;
                if ((rc = rd_sym(&d, &rnm, &rs))) { break; }
                if (rs == 63u) { break; }
                const uint32_t idx = pos + rs;
                if (idx > 63u) { rc = ORC_EMALFORMED; break; }
                const uint32_t lc = lvl_model == 1 ? pos_ctx(idx) : lvl_model == 2 ? prev_ctx(pcat, has_prev) : 0u;
                if ((rc = rd_sym(&d, &lvm[lc], &cat))) { break; }
                if ((rc = rd_suffix(&d, cat, &m))) { break; }
                if (m > lv_max) { rc = ORC_EMALFORMED; break; }
                blk[orc_zigzag[idx]] = unmap_signed(m);
                pcat = cat; has_prev = 1; pos = idx + 1u;
            }
            if (rc) { break; }
            crow[bx] = blk[0];
            store_block(po, blk, qm, qp, tbl, bd, bx, s->y0 + by);
        }
        int32_t* tmp = prow; prow = crow; crow = tmp;
    }
    free(prow);
    free(crow);
    if (rc == 0 && d.pos != d.size) { rc = ORC_EMALFORMED; }
    return rc;
}

/* Alpha slice: Rice pairs -> band-local MED reconstruction with [0,65535]
 * clamp and the mode-2 multiple check (native/src/codec/codec.c:7919-7972,
 * native/src/entropy/block_coding.c:529-561, slice_codec.c:1105-1127). */
static int decode_alpha(const orc_hdr* h, const orc_slice* s, const uint8_t* pl, const plane_out* po)
{
    const uint32_t w = h->pcw[3], rows = s->h * 8u;
    const size_t count = (size_t)w * rows;
    const uint32_t mask = h->alpha_mode == 2u ? (1u << (16u - h->alpha_bit_depth)) - 1u : 0u;
    uint16_t* band = (uint16_t*)malloc(count * 2);
    bitr br = { pl, s->size, 0 };
    size_t pos = 0;
    int rc = 0;
    while (pos < count && rc == 0) {
        const uint64_t remaining = count - pos;
        uint32_t rm, lm;
        if ((rc = rice_get(&br, s->k2, (uint32_t)remaining, &rm))) { break; }
        if ((rc = rice_get(&br, s->k1, 0x00020000u, &lm))) { break; }
        size_t end;
        int32_t level = 0;
        if (lm == 0u) {
            if ((uint64_t)rm != remaining) { rc = ORC_EMALFORMED; break; }
            end = pos + rm;
        } else {
            if ((uint64_t)rm + 1u > remaining) { rc = ORC_EMALFORMED; break; }
            level = unmap_signed(lm);
            end = pos + rm + 1u;
        }
        for (size_t i = pos; i < end; ++i) {
            const uint32_t y = (uint32_t)(i / w), x = (uint32_t)(i % w);
            const uint16_t* cur = band + (size_t)y * w;
            const uint16_t* up = y ? cur - w : cur;
            int32_t pred;
            if (y == 0) { pred = x == 0 ? 0 : cur[x - 1]; }
            else if (x == 0) { pred = up[0]; }
            else { pred = med(cur[x - 1], up[x], up[x - 1]); }
            int32_t v = pred + (lm != 0u && i == end - 1u ? level : 0);
            v = v < 0 ? 0 : v > 65535 ? 65535 : v;
            if (mask && ((uint32_t)v & mask)) { rc = ORC_EMALFORMED; break; }
            band[i] = (uint16_t)v;
        }
        pos = end;
        if (lm == 0u) { break; }
    }
    if (rc == 0) {
        const uint64_t aligned = (br.pos + 7u) & ~(uint64_t)7u;
        if (aligned > (uint64_t)s->size * 8u) { rc = ORC_ETRUNC; }
        else if (aligned != (uint64_t)s->size * 8u) { rc = ORC_EMALFORMED; }
    }
    if (rc == 0) {
        for (uint32_t y = 0; y < rows; ++y) {
            const uint32_t gy = s->y0 * 8u + y;
            if (gy >= po->vis_h) { break; }
            memcpy(po->dst + (size_t)gy * po->stride, band + (size_t)y * w, (size_t)po->vis_w * 2);
        }
    }
    free(band);
    return rc;
}

/* slice loop of a frame decode into planes of the source geometry: CRC,
 * entropy decode, reconstruction (coefficients up to scan position limit),
 * concealment of failed slices; alpha slices are skipped (status OK) when
 * skip_alpha (native/src/codec/codec.c:8206-8515, 8581-8644) */
static int32_t decode_core(const orc_hdr* hp, const uint8_t* pkt, size_t size, uint16_t* const planes[4],
                           const size_t strides[4], uint8_t* slice_status, uint32_t* nconc_out, uint32_t limit,
                           int skip_alpha)
{
    const orc_hdr h = *hp;
    const int bits = h.major == 1u || (h.major == 2u && h.entropy_mode == 1u);
    uint32_t nconc = 0;
    size_t off = 53;
    for (uint32_t i = 0; i < h.slice_count; ++i) {
        if (off + 17u > size) { return ORC_ETRUNC; }
        const uint8_t* sp = pkt + off;
        orc_slice s;
        s.size = ld32(sp); s.plane = sp[4]; s.y0 = (uint16_t)ld16(sp + 5); s.h = (uint16_t)ld16(sp + 7);
        s.qpd = sp[9]; s.k1 = sp[10]; s.k2 = sp[11]; s.k3 = sp[12]; s.crc = ld32(sp + 13);
        off += 17;
        if (off + s.size > size || s.plane >= h.plane_count) { return ORC_ETRUNC; }
        const uint8_t* pl = pkt + off;
        off += s.size;
        const uint32_t p = s.plane;
        if (skip_alpha && is_alpha(&h, p)) {
            if (slice_status) { slice_status[i] = 0u; }
            continue;
        }
        plane_out po = { planes[p], strides && strides[p] ? strides[p] : h.pvw[p], h.pvw[p], h.pvh[p], limit };
        int src = orc_crc32(pl, s.size) == s.crc ? 0 : ORC_ECRC;
        if (src == 0) {
            src = is_alpha(&h, p) ? decode_alpha(&h, &s, pl, &po)
                : bits ? decode_colour_bits(&h, &s, pl, &po) : decode_colour(&h, &s, pl, &po);
        }
        if (slice_status) { slice_status[i] = src ? 1u : 0u; }
        if (src) {
            nconc++;
            const uint16_t fill = is_alpha(&h, p) ? 0xFFFFu : (uint16_t)(1u << (h.bit_depth - 1u));
            const uint32_t y0 = s.y0 * 8u;
            uint32_t rows = s.h * 8u;
            if (rows > po.vis_h - y0) { rows = po.vis_h - y0; }
            for (uint32_t y = 0; y < rows; ++y) {
                for (uint32_t x = 0; x < po.vis_w; ++x) { po.dst[(size_t)(y0 + y) * po.stride + x] = fill; }
            }
        }
    }
    if (off != size) { return ORC_EMALFORMED; }
    *nconc_out = nconc;
    return ORC_OK;
}

static int32_t decode_header(const uint8_t* pkt, size_t size, orc_hdr* h)
{
    int rc = hdr_read(pkt, size, h);
    if (rc) { return rc; }
    const int bits = h->major == 1u || (h->major == 2u && h->entropy_mode == 1u);
    if (!(h->major == 7u && h->entropy_mode == 7u) && !bits) { return ORC_ENOTIMPL; }
    if (h->packet_size != size) { return ORC_EMALFORMED; }
    return ORC_OK;
}

int32_t orc_frame_decode(const uint8_t* pkt, size_t size, uint16_t* const planes[4],
                         const size_t strides[4], uint8_t* slice_status, uint32_t* concealed)
{
    return orc_frame_decode_ex(pkt, size, planes, strides, slice_status, concealed, 0);
}

/* full-size decode; skip_alpha leaves plane 3 out (decode request flag) */
int32_t orc_frame_decode_ex(const uint8_t* pkt, size_t size, uint16_t* const planes[4], const size_t strides[4],
                            uint8_t* slice_status, uint32_t* concealed, int32_t skip_alpha)
{
    orc_hdr h;
    int rc = decode_header(pkt, size, &h);
    if (rc) { return rc; }
    uint32_t nconc = 0;
    rc = decode_core(&h, pkt, size, planes, strides, slice_status, &nconc, 63u, skip_alpha && h.alpha_mode != 0u);
    if (rc) { return rc; }
    if (h.flags & 4u) {
        for (uint32_t p = 0; p < h.plane_count && p < 3u; ++p) {
            const size_t st = strides && strides[p] ? strides[p] : h.pvw[p];
            orc_deblock_plane(planes[p], st, h.pvw[p], h.pvh[p], h.qp_base, (h.flags & 8u) ? 1u : 0u,
                              h.bit_depth);
        }
    }
    if (concealed) { *concealed = nconc; }
    return nconc ? ORC_WARN_CONCEALED : ORC_OK;
}

/* Target-size decode (tc_frame_decode_scaled / _reduced / _request; native/
 * src/codec/codec.c:9100-9172, 9505-9640, 8581-8647 and color_store.h:
 * 104-288): the frame is reconstructed at source size keeping coefficients up
 * to scan position limit, then each target sample (tx, ty) of a plane takes
 * the source sample (floor(tx*vw/dw), floor(ty*vh/dh)); target plane sizes
 * are (tw, th), 4:2:2 chroma ((tw+1)/2, th), RAW phase planes ((tw+1)/2,
 * (th+1)/2). With the deblock flag and an exact 2:1 target (tw*2 = W and
 * th*2 = H) the colour target planes are filtered on a 4-sample grid
 * (codec.c:8649-8660, deblock.c:304-340). skip_alpha drops plane 3. */
int32_t orc_frame_decode_scaled(const uint8_t* pkt, size_t size, uint32_t tw, uint32_t th, uint32_t limit,
                                int32_t skip_alpha, uint16_t* const planes[4], const size_t strides[4],
                                uint8_t* slice_status, uint32_t* concealed)
{
    orc_hdr h;
    int rc = decode_header(pkt, size, &h);
    if (rc) { return rc; }
    if (tw == 0u || th == 0u || tw > h.visible_w || th > h.visible_h) { return ORC_EINVAL; }
    uint16_t* src[4] = { NULL, NULL, NULL, NULL };
    for (uint32_t p = 0; p < h.plane_count; ++p) {
        src[p] = (uint16_t*)calloc((size_t)h.pvw[p] * h.pvh[p], 2u);
        if (src[p] == NULL) { rc = ORC_ENOMEM; }
    }
    uint32_t nconc = 0;
    const int drop = skip_alpha && h.alpha_mode != 0u;
    if (rc == 0) { rc = decode_core(&h, pkt, size, src, NULL, slice_status, &nconc, limit, drop); }
    const int exact_half = tw * 2u == h.visible_w && th * 2u == h.visible_h;
    for (uint32_t p = 0; p < h.plane_count && rc == 0; ++p) {
        if (drop && p == 3u) { continue; }
        uint32_t dw = tw, dh = th;
        if (h.pixel_format == 3u) {
            dw = (tw + 1u) / 2u; dh = (th + 1u) / 2u;
        } else if (h.pixel_format == 0u && (p == 1u || p == 2u)) {
            dw = (tw + 1u) / 2u;
        }
        const size_t ds = strides && strides[p] ? strides[p] : dw;
        const uint32_t vw = h.pvw[p], vh = h.pvh[p];
        for (uint32_t ty = 0; ty < dh; ++ty) {
            const uint32_t sy = (uint32_t)(((uint64_t)ty * vh) / dh);
            for (uint32_t tx = 0; tx < dw; ++tx) {
                const uint32_t sx = (uint32_t)(((uint64_t)tx * vw) / dw);
                planes[p][(size_t)ty * ds + tx] = src[p][(size_t)sy * vw + sx];
            }
        }
        if ((h.flags & 4u) && p < 3u && exact_half) {
            orc_deblock_plane_unit(planes[p], ds, dw, dh, h.qp_base, (h.flags & 8u) ? 1u : 0u, h.bit_depth, 4u);
        }
    }
    for (uint32_t p = 0; p < 4u; ++p) { free(src[p]); }
    if (rc) { return rc; }
    if (concealed) { *concealed = nconc; }
    return nconc ? ORC_WARN_CONCEALED : ORC_OK;
}

/* -------------------------------------------------------- deblock */

static int32_t clampv(int32_t v, int32_t maxv) { return v < 0 ? 0 : v > maxv ? maxv : v; }
static uint32_t absu(int32_t v) { return (uint32_t)(v < 0 ? -v : v); }

static void db_edge(uint16_t* p1, uint16_t* p0, uint16_t* q0, uint16_t* q1,
                    int32_t alpha, int32_t beta, int32_t clip, int32_t maxv)
{
    const int32_t a = *p1, b = *p0, c = *q0, e = *q1;
    if (!(absu(c - b) < (uint32_t)alpha && absu(a - b) < (uint32_t)beta && absu(e - c) < (uint32_t)beta)) {
        return;
    }
    int32_t d = (12 * (c - b) + 4 * (a - e) + 8) >> 4;
    const int32_t delta = c - b;
    const int32_t half = delta >= 0 ? delta >> 1 : -((-delta) >> 1);
    if (half >= 0) { if (d > half) { d = half; } if (d < 0) { d = 0; } }
    else { if (d < half) { d = half; } if (d > 0) { d = 0; } }
    d = d < -clip ? -clip : d > clip ? clip : d;
    *p0 = (uint16_t)clampv(b + d, maxv);
    *q0 = (uint16_t)clampv(c - d, maxv);
    const int32_t ds = d >> 1;
    *p1 = (uint16_t)clampv(a + ds, maxv);
    *q1 = (uint16_t)clampv(e - ds, maxv);
}

/* Output weak filter on a grid of unit samples (8 for full-size output, 4 on
 * an exact 2:1 target): all vertical edges, then all horizontal edges
 * (native/src/codec/deblock.c:71-157,163-212,304-340). */
void orc_deblock_plane_unit(uint16_t* pix, size_t stride, uint32_t w, uint32_t h, uint32_t qp, uint32_t tbl,
                            uint32_t bd, uint32_t unit)
{
    if (pix == NULL || qp > 95u || unit < 4u || unit > 8u || w < 2u * unit || h < 2u * unit || (w % unit) ||
        (h % unit) || bd < 8u || bd > 16u) {
        return;
    }
    const int32_t maxv = (int32_t)((1u << bd) - 1u);
    uint32_t step = (orc_qp_scale[tbl & 1u][qp] * 3u + 173056u / 2u) / 173056u;
    if (step < 1u) { step = 1u; }
    uint32_t al = step * 2u;
    if (al < 3u) { al = 3u; }
    if (al > 48u) { al = 48u; }
    const int32_t alpha = (int32_t)al, beta = alpha / 2 + 2, clip = alpha >> 1;
    for (uint32_t y = 0; y < h; ++y) {
        uint16_t* r = pix + (size_t)y * stride;
        for (uint32_t x0 = unit; x0 + 1u < w; x0 += unit) {
            db_edge(r + x0 - 2, r + x0 - 1, r + x0, r + x0 + 1, alpha, beta, clip, maxv);
        }
    }
    for (uint32_t rr = 1; rr < h / unit; ++rr) {
        const uint32_t y0 = rr * unit;
        for (uint32_t x = 0; x < w; ++x) {
            db_edge(pix + (size_t)(y0 - 2) * stride + x, pix + (size_t)(y0 - 1) * stride + x,
                    pix + (size_t)y0 * stride + x, pix + (size_t)(y0 + 1) * stride + x,
                    alpha, beta, clip, maxv);
        }
    }
}

void orc_deblock_plane(uint16_t* pix, size_t stride, uint32_t w, uint32_t h, uint32_t qp, uint32_t tbl,
                       uint32_t bd)
{
    orc_deblock_plane_unit(pix, stride, w, h, qp, tbl, bd, 8u);
}

/* -------------------------------------------------- test image synthesiser */

/* xorshift64* driven test images of the reference test suite (flat, gradient,
 * grain, detail, mixed quadrants; 10-bit patterns x4 for 12-bit; 4:2:2 or
 * full-width chroma; alpha band pattern), restated from
 * native/tests/support/image_synth.c:1-150. Used only to regenerate the
 * inputs behind the reference's golden packets. */
static uint64_t isyn_next(uint64_t* s)
{
    uint64_t x = *s;
    x ^= x >> 12; x ^= x << 25; x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1Dull;
}

static uint32_t isyn_clamp(int32_t v, uint32_t scale)
{
    if (scale == 1u) { return v < 0 ? 0u : v > 1023 ? 1023u : (uint32_t)v; }
    int64_t s = (int64_t)v * scale;
    return s < 0 ? 0u : s > 4095 ? 4095u : (uint32_t)s;
}

static int32_t isyn_luma(uint32_t kind, uint32_t x, uint32_t y, uint32_t w, uint32_t h, uint64_t* s)
{
    uint32_t k = kind;
    if (kind == 4u) {
        const uint32_t q = (y >= h / 2u ? 2u : 0u) + (x >= w / 2u ? 1u : 0u);
        k = q == 0u ? 0u : q == 1u ? 1u : q == 2u ? 2u : 3u;
    }
    switch (k) {
    case 0: return (int32_t)(320u + (((x / 32u) + (y / 32u)) % 3u) * 6u);
    case 1: {
        const int64_t gx = (int64_t)x * 700 / (int64_t)(w > 1u ? w - 1u : 1u);
        const int64_t gy = (int64_t)y * 300 / (int64_t)(h > 1u ? h - 1u : 1u);
        return (int32_t)(64 + gx + gy);
    }
    case 2: {
        const uint64_t r = isyn_next(s);
        return 350 + (int32_t)(((x / 48u) + (y / 48u)) % 4u) * 40 + (int32_t)(r % 13u) - 6;
    }
    case 3: return (((x / 8u) + (y / 8u)) % 2u == 0u ? 180 : 760) + (((x + y) % 16u) < 2u ? 200 : 0);
    default: return 512;
    }
}

int orc_image_synth(uint64_t seed, uint32_t w, uint32_t h, uint32_t kind, uint32_t bit_depth,
                    uint32_t full_chroma, uint16_t* y, uint16_t* u, uint16_t* v, uint16_t* a)
{
    const uint32_t cw = full_chroma ? w : (w + 1u) / 2u;
    const uint32_t scale = bit_depth >= 12u ? 4u : 1u;
    uint64_t sy = seed ^ 0x4D4F565549564531ull, su = seed ^ 0x5543425732534544ull;
    uint64_t sv = seed ^ 0x5643425732534544ull, sa = seed ^ 0x414C504841323131ull;
    for (int i = 0; i < 8; ++i) { isyn_next(&sy); isyn_next(&su); isyn_next(&sv); isyn_next(&sa); }
    for (uint32_t yy = 0; yy < h; ++yy) {
        for (uint32_t x = 0; x < w; ++x) {
            const int32_t l = isyn_luma(kind, x, yy, w, h, &sy);
            if (y) { y[(size_t)yy * w + x] = (uint16_t)isyn_clamp(l, scale); }
        }
    }
    for (uint32_t yy = 0; yy < h; ++yy) {
        for (uint32_t x = 0; x < cw; ++x) {
            int32_t bu = 512 + (int32_t)((x / 24u) % 5u) * 24 - 48;
            int32_t bv = 512 + (int32_t)((yy / 24u) % 5u) * 20 - 40;
            if (kind == 2u || kind == 4u || kind == 3u) {
                bu += (int32_t)(isyn_next(&su) % 9u) - 4;
                bv += (int32_t)(isyn_next(&sv) % 9u) - 4;
            }
            if (u) { u[(size_t)yy * cw + x] = (uint16_t)isyn_clamp(bu, scale); }
            if (v) { v[(size_t)yy * cw + x] = (uint16_t)isyn_clamp(bv, scale); }
        }
    }
    if (a) {
        for (uint32_t yy = 0; yy < h; ++yy) {
            for (uint32_t x = 0; x < w; ++x) {
                uint32_t val;
                if (x < w / 4u) { val = 65535u; }
                else if (x < w / 2u) { val = (uint32_t)(65535ull * (x - w / 4u) / (w / 4u)); }
                else if (((x / 8u) + (yy / 8u)) % 2u == 0u) { val = 49152u; }
                else { val = 0u; }
                if (kind == 2u || kind == 4u) {
                    const uint32_t d = (uint32_t)(isyn_next(&sa) % 512u);
                    val = d > val ? 0u : val - d;
                }
                a[(size_t)yy * w + x] = (uint16_t)val;
            }
        }
    }
    return 0;
}

/* colour-plane 8x8 block count of a frame (sized encode's F-cache size rule) */
uint64_t orc_color_blocks(const orc_cfg* cfg)
{
    orc_hdr h;
    cfg_to_hdr(cfg, &h);
    derive_geometry(&h);
    const uint32_t colour = h.pixel_format == 3u ? 4u : 3u;
    uint64_t n = 0;
    for (uint32_t p = 0; p < colour; ++p) { n += (uint64_t)h.bcols[p] * h.brows[p]; }
    return n;
}
