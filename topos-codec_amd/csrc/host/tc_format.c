/* Frame format rules of the Topos intra bitstream on the host side: config
 * validation, config -> frame header, per-plane geometry, band plan, the
 * 53-byte frame header and 17-byte slice header codecs, the packet structure
 * parse, quantiser contexts and the frozen tables the format defines.
 *
 * Semantics follow (paths relative to the reference tree):
 *   config rules        native/src/codec/codec.c:426-605
 *   config -> header    native/src/codec/codec.c:280-397
 *   geometry            native/src/bitstream/frame_header.c:541-588
 *   header rules        native/src/bitstream/frame_header.c:99-534
 *   header codec        native/src/bitstream/frame_header.c:590-631, 703-781
 *   slice map           native/src/bitstream/slice_map.c:9-150
 *   structure parse     native/src/bitstream/packet.c:13-67
 *   quantiser context   native/src/transform/quant.c:217-308
 */
#include "tc_host.h"
#include "topos_gpu.h"
#include "../kernels/tc_math.h"

#include <string.h>

/* qp -> scale, classic and refined (QPT2) tables; quant matrices id 0..4.
 * Bitstream constants (native/src/transform/quant.c:18-196). */
const uint32_t tc_qp_scale_tab[2][96] = {
    { 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5, 6, 7,
      8, 10, 11, 13, 16, 19, 23, 27, 32, 38, 45, 54, 64, 76, 91, 108,
      128, 152, 181, 215, 256, 304, 362, 431, 512, 609, 724, 861, 1024, 1218, 1448, 1722,
      2048, 2435, 2896, 3444, 4096, 4871, 5793, 6889, 8192, 9742, 11585, 13777, 16384, 19484, 23170, 27555,
      32768, 38968, 46341, 55109, 65536, 77936, 92682, 110218, 131072, 155872, 185364, 220436,
      262144, 311744, 370728, 440872, 524288, 623487, 741455, 881744, 1048576, 1246974, 1482910, 1763488,
      2097152, 2493948, 2965821, 3526975, 4194304, 4987896, 5931642, 7053950 },
    { 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4, 5, 6, 7,
      8, 10, 11, 13, 16, 19, 23, 27, 32, 38, 45, 54, 64, 76, 91, 108,
      128, 152, 181, 215, 256, 304, 362, 431, 512, 609, 724, 861, 1024, 1218, 1448, 1722,
      2048, 2435, 2896, 3444, 4096, 4871, 5793, 6889, 8192, 9742, 11585, 13777, 16384, 19484, 23170, 27555,
      32768, 36781, 41285, 46341, 52016, 58386, 65536, 73562, 82570, 92682, 104032, 116772,
      131072, 147123, 165140, 185364, 208064, 233544, 262144, 294247, 330281, 370728, 416128, 467088,
      524288, 588493, 660561, 741455, 832255, 934175, 1048576, 1176987 },
};

#define R16 16, 16, 16, 16, 16, 16, 16, 16
const uint16_t tc_qm_tab[5][2][64] = {
    { { R16, R16, R16, R16, R16, R16, R16, R16 }, { R16, R16, R16, R16, R16, R16, R16, R16 } },
    { { 16, 20, 25, 31, 39, 49, 61, 76, 20, 25, 31, 39, 49, 61, 76, 95,
        25, 31, 39, 49, 61, 76, 95, 119, 31, 39, 49, 61, 76, 95, 119, 149,
        39, 49, 61, 76, 95, 119, 149, 186, 49, 61, 76, 95, 119, 149, 186, 233,
        61, 76, 95, 119, 149, 186, 233, 291, 76, 95, 119, 149, 186, 233, 291, 364 },
      { 16, 19, 22, 26, 31, 37, 43, 51, 19, 22, 26, 31, 37, 43, 51, 60,
        22, 26, 31, 37, 43, 51, 60, 71, 26, 31, 37, 43, 51, 60, 71, 84,
        31, 37, 43, 51, 60, 71, 84, 99, 37, 43, 51, 60, 71, 84, 99, 117,
        43, 51, 60, 71, 84, 99, 117, 138, 51, 60, 71, 84, 99, 117, 138, 162 } },
    { { 16, 36, 45, 56, 70, 88, 110, 137, 36, 45, 56, 70, 88, 110, 137, 171,
        45, 56, 70, 88, 110, 137, 171, 214, 56, 70, 88, 110, 137, 171, 214, 268,
        70, 88, 110, 137, 171, 214, 268, 335, 88, 110, 137, 171, 214, 268, 335, 419,
        110, 137, 171, 214, 268, 335, 419, 524, 137, 171, 214, 268, 335, 419, 524, 655 },
      { 16, 28, 33, 39, 46, 56, 64, 76, 28, 33, 39, 46, 56, 64, 76, 90,
        33, 39, 46, 56, 64, 76, 90, 106, 39, 46, 56, 64, 76, 90, 106, 126,
        46, 56, 64, 76, 90, 106, 126, 148, 56, 64, 76, 90, 106, 126, 148, 176,
        64, 76, 90, 106, 126, 148, 176, 207, 76, 90, 106, 126, 148, 176, 207, 243 } },
    { { 16, 22, 27, 33, 42, 53, 66, 82, 22, 27, 33, 42, 53, 66, 82, 103,
        27, 33, 42, 53, 66, 82, 103, 129, 33, 42, 53, 66, 82, 103, 129, 161,
        42, 53, 66, 82, 103, 129, 161, 201, 53, 66, 82, 103, 129, 161, 201, 252,
        66, 82, 103, 129, 161, 201, 252, 314, 82, 103, 129, 161, 201, 252, 314, 393 },
      { 16, 20, 23, 27, 33, 39, 45, 54, 20, 23, 27, 33, 39, 45, 54, 63,
        23, 27, 33, 39, 45, 54, 63, 75, 27, 33, 39, 45, 54, 63, 75, 88,
        33, 39, 45, 54, 63, 75, 88, 104, 39, 45, 54, 63, 75, 88, 104, 123,
        45, 54, 63, 75, 88, 104, 123, 145, 54, 63, 75, 88, 104, 123, 145, 170 } },
    { { 20, 19, 18, 17, 16, 15, 14, 13, 19, 18, 17, 16, 15, 14, 13, 12,
        18, 17, 16, 15, 14, 13, 12, 11, 17, 16, 15, 14, 13, 12, 11, 11,
        16, 15, 14, 13, 12, 11, 11, 10, 15, 14, 13, 12, 11, 11, 10, 9,
        14, 13, 12, 11, 11, 10, 9, 9, 13, 12, 11, 11, 10, 9, 9, 8 },
      { 18, 17, 16, 16, 15, 14, 13, 13, 17, 16, 16, 15, 14, 13, 13, 12,
        16, 16, 15, 14, 13, 13, 12, 11, 16, 15, 14, 13, 13, 12, 11, 11,
        15, 14, 13, 13, 12, 11, 11, 10, 14, 13, 13, 12, 11, 11, 10, 10,
        13, 13, 12, 11, 11, 10, 10, 9, 13, 12, 11, 11, 10, 10, 9, 9 } },
};
#undef R16

static void st16(uint8_t* p, uint32_t v) { p[0] = (uint8_t)(v >> 8); p[1] = (uint8_t)v; }
static void st32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)(v >> 24); p[1] = (uint8_t)(v >> 16); p[2] = (uint8_t)(v >> 8); p[3] = (uint8_t)v;
}
static uint32_t ld16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }
static uint32_t ld32(const uint8_t* p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}

/* ------------------------------------------------------------------ CRC32 */

static uint32_t g_crc_tab[256];
static int g_crc_ready = 0;

static void crc_init(void)
{
    if (g_crc_ready) { return; }
    for (uint32_t i = 0; i < 256u; ++i) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) { c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : (c >> 1); }
        g_crc_tab[i] = c;
    }
    g_crc_ready = 1;
}

uint32_t tc_crc32_update(uint32_t crc, const uint8_t* p, size_t n)
{
    crc_init();
    uint32_t c = crc ^ 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) { c = g_crc_tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8); }
    return c ^ 0xFFFFFFFFu;
}

uint32_t tc_crc32(const uint8_t* p, size_t n) { return tc_crc32_update(0u, p, n); }

void tc_crc_tables(uint32_t crc256[256], uint32_t x2n32[32])
{
    crc_init();
    memcpy(crc256, g_crc_tab, sizeof(g_crc_tab));
    /* x^(2^k) mod P in the reflected order: x^1 is bit 30 */
    uint32_t p = 0x40000000u;
    for (int k = 0; k < 32; ++k) {
        x2n32[k] = p;
        uint32_t sq = 0u, a = p, b = p;
        for (int i = 0; i < 32; ++i) {
            if (a & (0x80000000u >> i)) { sq ^= b; }
            b = (b & 1u) ? (b >> 1) ^ 0xEDB88320u : (b >> 1);
        }
        p = sq;
    }
}

/* log2 by exponent split and the atanh series up to z^13: the one floating
 * point routine of the encoder (its model choice compares exact f64 sums),
 * restated from native/src/codec/ctx_ceiling.c:75-89. Built with
 * -ffp-contract=off. */
double tc_ctx_log2(double x)
{
    uint64_t bits;
    memcpy(&bits, &x, 8);
    const int e = (int)((bits >> 52) & 0x7FFu) - 1023;
    bits = (bits & ~(0x7FFull << 52)) | (1023ull << 52);
    double m;
    memcpy(&m, &bits, 8);
    const double z = (m - 1.0) / (m + 1.0);
    const double z2 = z * z;
    const double at = z * (1.0 + z2 * (1.0 / 3.0 + z2 * (1.0 / 5.0 + z2 * (1.0 / 7.0
        + z2 * (1.0 / 9.0 + z2 * (1.0 / 11.0 + z2 * (1.0 / 13.0)))))));
    return (double)e + 2.885390081777926814 * at;
}

/* ------------------------------------------------------------ config rules */

static uint32_t pad8(uint32_t v) { return (v + 7u) / 8u * 8u; }

uint32_t tc_slice_rows_of(const topos_frame_config* cfg)
{
    return cfg->slice_rows != 0u ? cfg->slice_rows : TC_FRAME_DEFAULT_SLICE_ROWS;
}

static void select_minor(tc_fhdr* fh)
{
    if (fh->major >= 2u) { fh->minor = 0u; return; }
    fh->minor = fh->bit_depth == 16u ? 6u : fh->pixel_format == 3u ? 5u : fh->pixel_format == 2u ? 3u
              : fh->pixel_format == 1u ? 2u : fh->bit_depth != 10u ? 1u : 0u;
    if (fh->qp_base >= 64u && fh->minor < 4u) { fh->minor = 4u; }
    if (fh->color_transfer == 13u && fh->minor < 7u) { fh->minor = 7u; }
}

void tc_cfg_to_fhdr(const topos_frame_config* cfg, tc_fhdr* fh)
{
    memset(fh, 0, sizeof(*fh));
    const uint32_t em = cfg->reserved[0] & 0xFFu;
    fh->flags = cfg->alpha_premultiplied ? 1u : 0u;
    if (em == 8u && cfg->reserved[5] != 0u) {
        fh->flags = (uint16_t)(fh->flags | ((cfg->reserved[5] & 1u) ? 4u : 0u)
                               | ((cfg->reserved[5] & 2u) ? 8u : 0u));
    }
    fh->profile = cfg->profile ? cfg->profile : (cfg->pixel_format == 3u ? 7u : 3u);
    fh->pixel_format = cfg->pixel_format;
    fh->bit_depth = cfg->bit_depth ? cfg->bit_depth : 10u;
    fh->color_matrix = cfg->color_matrix ? cfg->color_matrix
                       : (fh->pixel_format == 2u || fh->pixel_format == 3u ? 0u : 1u);
    fh->qp_base = cfg->qp_base;
    select_minor(fh);
    fh->alpha_mode = cfg->alpha_mode;
    fh->alpha_bit_depth = cfg->alpha_bit_depth;
    fh->coded_w = (uint16_t)pad8(cfg->visible_width);
    fh->coded_h = (uint16_t)pad8(cfg->visible_height);
    fh->visible_w = cfg->visible_width;
    fh->visible_h = cfg->visible_height;
    fh->plane_count = (uint8_t)(cfg->pixel_format == 3u ? 4u : 3u + (cfg->alpha_mode ? 1u : 0u));
    fh->qmatrix_id = cfg->qmatrix_id;
    fh->color_range = cfg->color_range;
    fh->color_primaries = cfg->color_primaries ? cfg->color_primaries : 1u;
    fh->color_transfer = cfg->color_transfer ? cfg->color_transfer
                         : (cfg->pixel_format == 3u ? (cfg->bit_depth == 16u ? 8u : TC_TRANSFER_TRAW_LOG)
                                                    : 13u);
    fh->chroma_siting = cfg->chroma_siting;
    select_minor(fh);
    fh->sar_num = cfg->sar_num;
    fh->sar_den = cfg->sar_den;
    if (em == 8u) { fh->major = 7u; fh->entropy_mode = 7u; }
    else if (em == 9u) { fh->major = 8u; fh->entropy_mode = 8u; }
    else if (em == 10u) {
        fh->major = 9u; fh->entropy_mode = 9u;
        fh->frame_type = (uint8_t)cfg->reserved[3];
        fh->gop_id = (uint16_t)cfg->reserved[4];
        fh->ref_distance = (uint8_t)cfg->reserved[5];
    } else if (em == 11u) {
        fh->major = 7u; fh->entropy_mode = 8u;
        fh->frame_type = (uint8_t)cfg->reserved[3];
        fh->gop_id = (uint16_t)cfg->reserved[4];
        fh->ref_distance = (uint8_t)cfg->reserved[5];
    } else if (em == 1u) { fh->major = 2u; fh->entropy_mode = 1u; fh->codebook_version = 1u; }
    else { fh->major = 1u; }
    if (fh->major >= 2u) { select_minor(fh); }
}

int32_t tc_fhdr_geometry(tc_fhdr* fh)
{
    if (fh->pixel_format == 3u) {
        const uint32_t vw = ((uint32_t)fh->visible_w + 1u) / 2u, vh = ((uint32_t)fh->visible_h + 1u) / 2u;
        const uint32_t cw = pad8(vw), ch = pad8(vh);
        for (int p = 0; p < 4; ++p) {
            fh->pcw[p] = (uint16_t)cw; fh->pch[p] = (uint16_t)ch;
            fh->pvw[p] = (uint16_t)vw; fh->pvh[p] = (uint16_t)vh;
            fh->bcols[p] = (uint16_t)(cw / 8u); fh->brows[p] = (uint16_t)(ch / 8u);
        }
        if (cw > TC_MAX_DIM || ch > TC_MAX_DIM) {
            return tc_fail(TC_ERR_LIMIT_EXCEEDED, "cfa plane coded dim exceeds 16384");
        }
        return TC_OK;
    }
    const uint32_t cvw = fh->pixel_format == 0u ? ((uint32_t)fh->visible_w + 1u) / 2u : fh->visible_w;
    const uint32_t ccw = fh->pixel_format == 0u ? pad8(cvw) : fh->coded_w;
    for (int p = 0; p < 4; ++p) {
        const int ch = p == 1 || p == 2;
        fh->pcw[p] = (uint16_t)(ch ? ccw : fh->coded_w);
        fh->pch[p] = fh->coded_h;
        fh->pvw[p] = (uint16_t)(ch ? cvw : fh->visible_w);
        fh->pvh[p] = fh->visible_h;
        fh->bcols[p] = (uint16_t)(fh->pcw[p] / 8u);
        fh->brows[p] = (uint16_t)(fh->pch[p] / 8u);
    }
    if (fh->pcw[1] > TC_MAX_DIM || fh->pch[1] > TC_MAX_DIM) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "chroma coded dim exceeds 16384");
    }
    return TC_OK;
}

int32_t tc_band_plan(const tc_fhdr* fh, uint32_t slice_rows, uint32_t bands[4], uint32_t* total)
{
    uint64_t t = 0u;
    for (uint32_t p = 0; p < fh->plane_count; ++p) {
        uint64_t b = ((uint64_t)fh->brows[p] + slice_rows - 1u) / slice_rows;
        if (b == 0u) { b = 1u; }
        if (bands) { bands[p] = (uint32_t)b; }
        t += b;
    }
    if (t > TC_MAX_SLICE_COUNT) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "slice_rows %u yields %llu slices > 512",
                       (unsigned)slice_rows, (unsigned long long)t);
    }
    if (total) { *total = (uint32_t)t; }
    return TC_OK;
}

static int32_t vfail(int32_t code, const char* field, unsigned long long v)
{
    return tc_fail(code, "frame header field %s = %llu rejected", field, v);
}

/* Header field rules, checked in the reference's order so the first failing
 * rule yields the same status (native/src/bitstream/frame_header.c:99-534). */
int32_t tc_fhdr_validate(const tc_fhdr* fh)
{
    const uint32_t mj = fh->major;
    if (mj >= 2u && mj <= 9u) {
        if (mj >= 3u && mj <= 6u) { return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "generation %u retired", mj); }
        if (mj == 2u && fh->entropy_mode == 0u) { return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "generation 2 retired"); }
        if (mj == 7u && fh->entropy_mode == 6u) { return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "generation 7 retired"); }
        if (fh->minor != 0u) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "v2 version_minor", fh->minor); }
        if (mj == 7u) {
            if (!((fh->entropy_mode == 6u || fh->entropy_mode == 7u || fh->entropy_mode == 8u) &&
                  fh->codebook_version == 0u && fh->coding_mode == 0u)) {
                return vfail(TC_ERR_UNSUPPORTED_VERSION, "v7 rans entropy fields", fh->entropy_mode);
            }
        } else if (mj == 8u) {
            if (!(fh->entropy_mode == 8u && fh->codebook_version == 0u && fh->coding_mode == 0u)) {
                return vfail(TC_ERR_UNSUPPORTED_VERSION, "v8 entropy fields", fh->entropy_mode);
            }
        } else if (mj == 9u) {
            if (!(fh->entropy_mode == 9u && fh->codebook_version == 0u && fh->coding_mode == 0u)) {
                return vfail(TC_ERR_UNSUPPORTED_VERSION, "v9 entropy fields", fh->entropy_mode);
            }
        } else if (fh->entropy_mode > 1u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "entropy_mode", fh->entropy_mode);
        }
        if (mj < 7u && (fh->entropy_mode == 0u ? fh->codebook_version != 0u : fh->codebook_version != 1u)) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "codebook_version", fh->codebook_version);
        }
        if (mj < 7u && fh->coding_mode != 0u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "coding_mode", fh->coding_mode);
        }
        if ((fh->flags & ~(uint16_t)0x0003u) != 0u) {
            const uint16_t nb = (uint16_t)(fh->flags & 0x000Cu);
            if ((fh->flags & ~(uint16_t)0x000Fu) != 0u) {
                return vfail(TC_ERR_MALFORMED, "v2 flags(reserved bits)", fh->flags);
            }
            if (nb != 0u) {
                const int dom = mj == 7u && fh->entropy_mode == 7u && fh->frame_type == 0u &&
                                fh->profile == 3u && fh->bit_depth == 10u;
                if (!dom) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "flags r4/r5 domain", fh->flags); }
            }
        }
        if ((fh->flags & 0x0002u) != 0u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "tile_layout(V2.0)", fh->flags);
        }
        if (mj == 9u || (mj == 7u && fh->entropy_mode == 8u)) {
            if (fh->frame_type > 1u) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "v9 frame_type", fh->frame_type); }
            if (fh->frame_type == 0u && fh->ref_distance != 0u) {
                return vfail(TC_ERR_MALFORMED, "v9 I ref_distance(!=0)", fh->ref_distance);
            }
            if (fh->frame_type == 1u && fh->ref_distance != 1u) {
                return vfail(TC_ERR_MALFORMED, "v9 P ref_distance(!=1)", fh->ref_distance);
            }
        } else if (fh->frame_type != 0u || fh->gop_id != 0u || fh->ref_distance != 0u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "frame_type/gop/ref(V2.1)", fh->frame_type);
        }
    } else if (mj != 1u) {
        return vfail(TC_ERR_UNSUPPORTED_VERSION, "version_major", mj);
    } else {
        if (fh->entropy_mode != 0u || fh->codebook_version != 0u || fh->coding_mode != 0u) {
            return vfail(TC_ERR_MALFORMED, "v1 entropy fields(!=0)", fh->entropy_mode);
        }
        if ((fh->flags & ~(uint16_t)0x0001u) != 0u) { return vfail(TC_ERR_MALFORMED, "flags(reserved bits)", fh->flags); }
        if (fh->frame_type != 0u) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "frame_type(V2.1)", fh->frame_type); }
        if (fh->gop_id != 0u) { return vfail(TC_ERR_MALFORMED, "gop_id(reserved)", fh->gop_id); }
        if (fh->ref_distance != 0u) { return vfail(TC_ERR_MALFORMED, "ref_distance(reserved)", fh->ref_distance); }
    }
    if (fh->profile < 1u || fh->profile > 7u) { return vfail(TC_ERR_UNSUPPORTED_PROFILE, "profile", fh->profile); }
    if (fh->profile == 7u) {
        if (fh->pixel_format != 3u) { return vfail(TC_ERR_MALFORMED, "profile(TRAW requires CFA pf=3)", fh->pixel_format); }
        if (fh->bit_depth != 12u && fh->bit_depth != 16u) {
            return vfail(TC_ERR_MALFORMED, "profile(TRAW requires 12/16-bit)", fh->bit_depth);
        }
        if (fh->alpha_mode != 0u) { return vfail(TC_ERR_MALFORMED, "profile(TRAW alpha unsupported)", fh->alpha_mode); }
        if (fh->qmatrix_id != 0u) { return vfail(TC_ERR_MALFORMED, "profile(TRAW frozen qm0)", fh->qmatrix_id); }
        const uint8_t tt = fh->bit_depth == 12u ? TC_TRANSFER_TRAW_LOG : 8u;
        if (fh->color_transfer != tt) {
            return vfail(TC_ERR_MALFORMED, "profile(TRAW transfer mismatch for bit depth)", fh->color_transfer);
        }
    } else {
        if (fh->profile != 3u && fh->profile != 5u && fh->profile != 6u) {
            return vfail(TC_ERR_UNSUPPORTED_PROFILE, "profile(not 3/5/6)", fh->profile);
        }
        if (fh->pixel_format == 3u) {
            return vfail(TC_ERR_MALFORMED, "pixel_format(CFA requires TRAW profile 7)", fh->pixel_format);
        }
    }
    if (fh->profile == 5u || fh->profile == 6u) {
        if (fh->pixel_format == 0u) {
            return vfail(TC_ERR_MALFORMED, "profile/format(Pro444|Extreme requires 4:4:4)", fh->pixel_format);
        }
        if (fh->profile == 6u && fh->bit_depth != 12u) {
            return vfail(TC_ERR_MALFORMED, "profile/format(Extreme requires 12-bit)", fh->bit_depth);
        }
    }
    if (fh->pixel_format > 3u) { return vfail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "pixel_format", fh->pixel_format); }
    if (fh->bit_depth != 10u && fh->bit_depth != 12u && fh->bit_depth != 16u) {
        return vfail(TC_ERR_UNSUPPORTED_PIXEL_FORMAT, "bit_depth", fh->bit_depth);
    }
    if (mj == 1u) {
        if (fh->minor > 7u) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "version_minor", fh->minor); }
        if (fh->minor == 0u && (fh->bit_depth != 10u || fh->pixel_format != 0u)) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.2 extended enum in version_minor=0 stream", fh->bit_depth);
        }
        if (fh->minor == 1u && fh->pixel_format != 0u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.3 extended enum in version_minor=1 stream", fh->pixel_format);
        }
        if (fh->minor == 2u && fh->pixel_format > 1u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.4 extended enum in version_minor=2 stream", fh->pixel_format);
        }
        if (fh->minor < 5u && (fh->pixel_format == 3u || fh->profile == 7u)) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.6 extended enum in version_minor<5 stream", fh->pixel_format);
        }
        if (fh->minor < 6u && fh->bit_depth == 16u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.7 extended enum in version_minor<6 stream", fh->bit_depth);
        }
    }
    if (fh->alpha_mode > 2u) { return vfail(TC_ERR_MALFORMED, "alpha_mode", fh->alpha_mode); }
    if (fh->alpha_mode == 0u) {
        if (fh->alpha_bit_depth != 0u) { return vfail(TC_ERR_MALFORMED, "alpha_bit_depth(!=0 w/o alpha)", fh->alpha_bit_depth); }
    } else if (fh->alpha_mode == 1u) {
        if (fh->alpha_bit_depth != 16u) { return vfail(TC_ERR_MALFORMED, "alpha_bit_depth(!=16 lossless)", fh->alpha_bit_depth); }
    } else if (fh->alpha_bit_depth != 8u && fh->alpha_bit_depth != 10u && fh->alpha_bit_depth != 12u) {
        return vfail(TC_ERR_MALFORMED, "alpha_bit_depth(near-lossless)", fh->alpha_bit_depth);
    }
    if (fh->coded_w == 0u || fh->coded_h == 0u || (fh->coded_w % 8u) != 0u || (fh->coded_h % 8u) != 0u ||
        fh->coded_w > TC_MAX_DIM || fh->coded_h > TC_MAX_DIM) {
        return vfail(TC_ERR_LIMIT_EXCEEDED, "coded_w/h(16384,%8)", fh->coded_w);
    }
    if (fh->visible_w == 0u || fh->visible_h == 0u) { return vfail(TC_ERR_MALFORMED, "visible_w/h(>=1)", fh->visible_w); }
    if (fh->visible_w > fh->coded_w || fh->visible_h > fh->coded_h) {
        return vfail(TC_ERR_LIMIT_EXCEEDED, "visible>coded", fh->visible_w);
    }
    const uint8_t ep = fh->pixel_format == 3u ? 4u : (uint8_t)(3u + (fh->alpha_mode ? 1u : 0u));
    if (fh->plane_count != ep) { return vfail(TC_ERR_MALFORMED, "plane_count", fh->plane_count); }
    if (fh->qmatrix_id > 4u) { return vfail(TC_ERR_UNSUPPORTED_MATRIX, "qmatrix_id", fh->qmatrix_id); }
    if (fh->qmatrix_id == 2u && !(fh->pixel_format == 2u && fh->bit_depth == 12u && (fh->profile == 5u || fh->profile == 6u))) {
        return vfail(TC_ERR_UNSUPPORTED_MATRIX, "qmatrix_id=2 requires GBR 4:4:4 12-bit profile 5/6", fh->qmatrix_id);
    }
    if ((fh->qmatrix_id == 3u || fh->qmatrix_id == 4u) &&
        !(fh->pixel_format == 0u && fh->bit_depth == 10u && fh->profile == 3u)) {
        return vfail(TC_ERR_UNSUPPORTED_MATRIX, "qmatrix_id=3/4 requires YUV 4:2:2 10-bit profile 3", fh->qmatrix_id);
    }
    if (fh->qp_base > TC_QP_MAX) { return vfail(TC_ERR_MALFORMED, "qp_base", fh->qp_base); }
    if (fh->qp_base >= 64u && !((mj == 1u && fh->minor >= 4u) || mj >= 2u)) {
        return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.5 qp>=64 in version_minor<4 stream", fh->qp_base);
    }
    if (fh->slice_count == 0u) { return vfail(TC_ERR_MALFORMED, "slice_count(>=1)", fh->slice_count); }
    if (fh->slice_count > TC_MAX_SLICE_COUNT) { return vfail(TC_ERR_LIMIT_EXCEEDED, "slice_count(<=512)", fh->slice_count); }
    if (fh->slice_count < fh->plane_count) { return vfail(TC_ERR_MALFORMED, "slice_count(<plane_count)", fh->slice_count); }
    if (fh->color_range > 1u) { return vfail(TC_ERR_MALFORMED, "color_range", fh->color_range); }
    if (fh->chroma_siting > 2u) { return vfail(TC_ERR_MALFORMED, "chroma_siting", fh->chroma_siting); }
    if ((fh->pixel_format == 2u || fh->pixel_format == 3u) && fh->chroma_siting != 0u) {
        return vfail(TC_ERR_MALFORMED, "chroma_siting(!=0 w/ GBR|CFA)", fh->chroma_siting);
    }
    switch (fh->color_primaries) {
    case 1: case 6: case 9: case 12: break;
    default: return vfail(TC_ERR_MALFORMED, "color_primaries(tag)", fh->color_primaries);
    }
    switch (fh->color_transfer) {
    case 1: case 8: case 16: case 18: case TC_TRANSFER_TRAW_LOG: break;
    case 13:
        if (mj == 1u && fh->minor < 7u) {
            return vfail(TC_ERR_UNSUPPORTED_VERSION, "v1.8 sRGB transfer(13) in version_minor<7 stream", fh->minor);
        }
        break;
    default: return vfail(TC_ERR_MALFORMED, "color_transfer(tag)", fh->color_transfer);
    }
    if (fh->profile == 7u) {
        const uint32_t tt = fh->bit_depth == 16u ? 8u : TC_TRANSFER_TRAW_LOG;
        if (fh->color_transfer != tt) { return vfail(TC_ERR_MALFORMED, "color_transfer(profile7 pair)", fh->color_transfer); }
    } else if (fh->color_transfer == TC_TRANSFER_TRAW_LOG) {
        return vfail(TC_ERR_MALFORMED, "color_transfer(TRAW_LOG0 is profile-7 private)", fh->color_transfer);
    }
    if (fh->pixel_format == 2u || fh->pixel_format == 3u) {
        if (fh->color_matrix != 0u) { return vfail(TC_ERR_MALFORMED, "color_matrix(!=0 identity for GBR)", fh->color_matrix); }
    } else {
        switch (fh->color_matrix) {
        case 1: case 5: case 9: break;
        default: return vfail(TC_ERR_MALFORMED, "color_matrix(tag)", fh->color_matrix);
        }
    }
    if ((fh->sar_num == 0u) != (fh->sar_den == 0u)) { return vfail(TC_ERR_MALFORMED, "sar(0 iff 0)", fh->sar_num); }
    if (fh->packet_size == 0u || fh->packet_size > TC_MAX_PACKET) {
        return vfail(TC_ERR_LIMIT_EXCEEDED, "frame_packet_size(<=256MiB)", fh->packet_size);
    }
    return TC_OK;
}

int32_t tc_cfg_validate(const topos_frame_config* cfg)
{
    if (cfg == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "cfg == NULL"); }
    if (cfg->struct_size != 0u && cfg->struct_size != (uint32_t)sizeof(topos_frame_config)) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "cfg struct_size %u != %u", (unsigned)cfg->struct_size,
                       (unsigned)sizeof(topos_frame_config));
    }
    if (cfg->abi_version != 0u && cfg->abi_version != (uint32_t)TOPOS_CODEC_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "cfg abi_version %u unsupported", (unsigned)cfg->abi_version);
    }
    if (cfg->visible_width == 0u || cfg->visible_height == 0u || cfg->visible_width > TC_MAX_DIM ||
        cfg->visible_height > TC_MAX_DIM) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "visible %ux%u out of 1..16384", (unsigned)cfg->visible_width,
                       (unsigned)cfg->visible_height);
    }
    if (cfg->qp_base > TC_QP_MAX) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "qp_base %u > 95", (unsigned)cfg->qp_base); }
    if (cfg->reserved[0] > 11u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "reserved[0] coding selection %u not in {0..11}", (unsigned)cfg->reserved[0]);
    }
    if (cfg->reserved[0] >= 2u && cfg->reserved[0] <= 7u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "reserved[0] coding selection %u retired", (unsigned)cfg->reserved[0]);
    }
    if (cfg->reserved[1] > 1u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "reserved[1] AQ mode %u not in {0,1}", (unsigned)cfg->reserved[1]); }
    if (cfg->reserved[2] > 1u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "reserved[2] RDO mode %u not in {0,1}", (unsigned)cfg->reserved[2]); }
    {
        const uint32_t em = cfg->reserved[0] & 0xFFu;
        const int rsv_end = em >= 10u ? 6 : (em == 9u ? 5 : 3);
        for (int i = rsv_end; i < 8; ++i) {
            if (em == 8u && i == 5) { continue; }
            if (cfg->reserved[i] != 0u) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT, "reserved[%d] must be 0 (got %u)", i, (unsigned)cfg->reserved[i]);
            }
        }
        if (em == 8u && cfg->reserved[5] != 0u) {
            const uint32_t f = cfg->reserved[5];
            if ((f & ~3u) != 0u || cfg->profile != 3u || cfg->pixel_format != 0u || cfg->bit_depth != 10u) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT,
                               "reserved[5] output features %u require profile 3 / 4:2:2 / 10-bit", (unsigned)f);
            }
        }
        if (em == 10u || em == 11u) {
            if (cfg->reserved[3] > 1u) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT, "temporal frame_type %u not in {0,1}", (unsigned)cfg->reserved[3]);
            }
            const uint32_t want = cfg->reserved[3] == 0u ? 0u : 1u;
            if (cfg->reserved[5] != want) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT, "temporal ref_distance %u != %u", (unsigned)cfg->reserved[5],
                               (unsigned)want);
            }
        }
    }
    if (cfg->alpha_mode > 2u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "alpha_mode %u > 2", (unsigned)cfg->alpha_mode); }
    if (cfg->alpha_mode == 1u && cfg->alpha_bit_depth != 16u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "lossless alpha requires bit_depth 16");
    }
    if (cfg->alpha_mode == 2u && cfg->alpha_bit_depth != 8u && cfg->alpha_bit_depth != 10u &&
        cfg->alpha_bit_depth != 12u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "near-lossless alpha bit_depth %u not in {8,10,12}",
                       (unsigned)cfg->alpha_bit_depth);
    }
    if (cfg->pixel_format == 3u) {
        if (cfg->alpha_mode != 0u) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "TRAW (pixel_format=3) does not support alpha"); }
        if (cfg->qp_delta_chroma != 0) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "qp_delta_chroma must be 0 for TRAW"); }
    }
    tc_fhdr fh;
    tc_cfg_to_fhdr(cfg, &fh);
    int32_t rc = tc_fhdr_geometry(&fh);
    if (rc != TC_OK) { return rc; }
    rc = tc_band_plan(&fh, tc_slice_rows_of(cfg), NULL, NULL);
    if (rc != TC_OK) { return rc; }
    fh.slice_count = (uint16_t)(cfg->pixel_format == 3u ? 4u : 3u + (cfg->alpha_mode ? 1u : 0u));
    fh.packet_size = TC_HDR_SIZE;
    rc = tc_fhdr_validate(&fh);
    if (rc != TC_OK) { return rc; }
    if (cfg->bit_depth >= 13u && cfg->reserved[0] != 8u && cfg->reserved[0] != 9u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "entropy mode %u is 12-bit frozen; bit_depth %u requires rans2",
                       (unsigned)cfg->reserved[0], (unsigned)cfg->bit_depth);
    }
    return TC_OK;
}

/* ------------------------------------------------------------ header codec */

void tc_fhdr_pack(const tc_fhdr* fh, uint8_t o[TC_HDR_SIZE])
{
    o[0] = 'T'; o[1] = 'P'; o[2] = 'I'; o[3] = 'C';
    st16(o + 4, TC_HDR_SIZE);
    o[6] = fh->major; o[7] = fh->minor; st16(o + 8, fh->flags);
    o[10] = fh->profile; o[11] = fh->pixel_format; o[12] = fh->bit_depth; o[13] = fh->alpha_mode;
    o[14] = fh->alpha_bit_depth; o[15] = fh->frame_type; st16(o + 16, fh->gop_id); o[18] = fh->ref_distance;
    st16(o + 19, fh->coded_w); st16(o + 21, fh->coded_h); st16(o + 23, fh->visible_w); st16(o + 25, fh->visible_h);
    o[27] = fh->plane_count; o[28] = fh->qmatrix_id; o[29] = fh->qp_base; st16(o + 30, fh->slice_count);
    o[32] = fh->color_range; o[33] = fh->color_primaries; o[34] = fh->color_transfer;
    o[35] = fh->color_matrix; o[36] = fh->chroma_siting; st16(o + 37, fh->sar_num); st16(o + 39, fh->sar_den);
    st32(o + 41, fh->packet_size);
    o[45] = fh->entropy_mode; o[46] = fh->codebook_version; o[47] = fh->coding_mode; o[48] = 0u;
    st32(o + 49, tc_crc32(o, 49u));
}

int32_t tc_fhdr_parse(const uint8_t* d, size_t size, tc_fhdr* fh)
{
    if (size < TC_HDR_SIZE) {
        return tc_fail(TC_ERR_TRUNCATED, "frame header needs 53 bytes, got %zu", size);
    }
    if (memcmp(d, "TPIC", 4) != 0) { return tc_fail(TC_ERR_MALFORMED, "magic mismatch (not TPIC)"); }
    const uint32_t crc = tc_crc32(d, 49u);
    if (ld32(d + 49) != crc) {
        return tc_fail(TC_ERR_CHECKSUM_MISMATCH, "frame header crc32 %08x != %08x", (unsigned)ld32(d + 49), (unsigned)crc);
    }
    memset(fh, 0, sizeof(*fh));
    if (ld16(d + 4) != TC_HDR_SIZE) { return vfail(TC_ERR_MALFORMED, "header_size(!=53)", ld16(d + 4)); }
    const uint8_t mj = d[6];
    if (mj >= 3u && mj <= 6u) { return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "generation %u retired", (unsigned)mj); }
    if (mj < 1u || mj > 9u) { return vfail(TC_ERR_UNSUPPORTED_VERSION, "version_major", mj); }
    fh->major = mj; fh->minor = d[7]; fh->flags = (uint16_t)ld16(d + 8);
    fh->profile = d[10]; fh->pixel_format = d[11]; fh->bit_depth = d[12]; fh->alpha_mode = d[13];
    fh->alpha_bit_depth = d[14]; fh->frame_type = d[15]; fh->gop_id = (uint16_t)ld16(d + 16);
    fh->ref_distance = d[18]; fh->coded_w = (uint16_t)ld16(d + 19); fh->coded_h = (uint16_t)ld16(d + 21);
    fh->visible_w = (uint16_t)ld16(d + 23); fh->visible_h = (uint16_t)ld16(d + 25);
    fh->plane_count = d[27]; fh->qmatrix_id = d[28]; fh->qp_base = d[29]; fh->slice_count = (uint16_t)ld16(d + 30);
    fh->color_range = d[32]; fh->color_primaries = d[33]; fh->color_transfer = d[34];
    fh->color_matrix = d[35]; fh->chroma_siting = d[36]; fh->sar_num = (uint16_t)ld16(d + 37);
    fh->sar_den = (uint16_t)ld16(d + 39); fh->packet_size = ld32(d + 41);
    if (mj == 1u) {
        if (ld32(d + 45) != 0u) { return vfail(TC_ERR_MALFORMED, "reserved0(!=0)", ld32(d + 45)); }
    } else {
        fh->entropy_mode = d[45]; fh->codebook_version = d[46]; fh->coding_mode = d[47];
        if (d[48] != 0u) { return vfail(TC_ERR_MALFORMED, "reserved_v2_0(!=0)", d[48]); }
    }
    int32_t rc = tc_fhdr_validate(fh);
    if (rc != TC_OK) { return rc; }
    return tc_fhdr_geometry(fh);
}

/* slice headers + exact exhaustion + plane/band tiling + per-mode k rules
 * (native/src/bitstream/packet.c:13-67, slice_map.c:23-150) */
int32_t tc_parse_structure(const uint8_t* data, size_t size, tc_view* v)
{
    memset(&v->fh, 0, sizeof(v->fh));
    v->nslices = 0;
    int32_t rc = tc_fhdr_parse(data, size, &v->fh);
    if (rc != TC_OK) { return rc; }
    const tc_fhdr* fh = &v->fh;
    if (fh->major == 8u || fh->major == 9u) {
        return tc_fail(TC_ERR_NOT_IMPLEMENTED, "segmented (major %u) packets are not decoded by this backend",
                       (unsigned)fh->major);
    }
    if ((uint64_t)fh->packet_size != (uint64_t)size) {
        return tc_fail(TC_ERR_MALFORMED, "frame_packet_size %u != actual %zu", (unsigned)fh->packet_size, size);
    }
    const uint16_t n = fh->slice_count;
    size_t off = TC_HDR_SIZE;
    for (uint16_t i = 0; i < n; ++i) {
        if (off > size || size - off < TC_SLICE_HDR_SIZE) {
            return tc_fail(TC_ERR_TRUNCATED, "slice %u header out of packet at offset %zu", (unsigned)i, off);
        }
        const uint8_t* s = data + off;
        tc_shdr* h = &v->sl[i];
        h->size = ld32(s); h->plane = s[4]; h->y0 = (uint16_t)ld16(s + 5); h->h = (uint16_t)ld16(s + 7);
        h->qpd = s[9]; h->k1 = s[10]; h->k2 = s[11]; h->k3 = s[12]; h->crc = ld32(s + 13);
        if (h->k1 > 14u || h->k2 > 14u || h->k3 > 14u) {
            return tc_fail(TC_ERR_MALFORMED, "slice k1/k2/k3 > 14: %u/%u/%u", h->k1, h->k2, h->k3);
        }
        off += TC_SLICE_HDR_SIZE;
        if (size - off < (size_t)h->size) {
            return tc_fail(TC_ERR_TRUNCATED, "slice %u payload %u out of packet at offset %zu", (unsigned)i,
                           (unsigned)h->size, off);
        }
        v->off[i] = (uint32_t)off;
        off += h->size;
    }
    if (off != size) { return tc_fail(TC_ERR_MALFORMED, "packet has %zu trailing bytes after last slice", size - off); }
    v->nslices = n;
    uint32_t idx = 0u;
    for (uint32_t p = 0; p < fh->plane_count; ++p) {
        uint32_t covered = 0u;
        const uint32_t rows = fh->brows[p];
        while (covered < rows) {
            if (idx >= n) { return tc_fail(TC_ERR_MALFORMED, "plane %u tiling incomplete at block row %u", p, covered); }
            const tc_shdr* h = &v->sl[idx];
            if (h->plane != p) { return tc_fail(TC_ERR_MALFORMED, "slice %u plane %u breaks Y->U->V->[A] order", idx, h->plane); }
            const int colour = p < 3u || fh->alpha_mode == 0u;
            if (fh->major >= 2u && fh->entropy_mode == 1u && colour && (h->k1 > 3u || h->k2 > 3u || h->k3 > 3u)) {
                return tc_fail(TC_ERR_MALFORMED, "slice %u vlc book id > 3", idx);
            }
            if (fh->major == 7u && (fh->entropy_mode == 6u || fh->entropy_mode == 7u || fh->entropy_mode == 8u) &&
                colour && (h->k1 != 0u || h->k2 != 0u || h->k3 != 0u)) {
                return tc_fail(TC_ERR_MALFORMED, "slice %u rans k1/k2/k3 != 0", idx);
            }
            if (h->y0 != covered) { return tc_fail(TC_ERR_MALFORMED, "slice %u block_y0 %u != expected %u", idx, h->y0, covered); }
            if (h->h == 0u) { return tc_fail(TC_ERR_MALFORMED, "slice %u block_h == 0", idx); }
            if ((uint32_t)h->y0 + h->h > rows) { return tc_fail(TC_ERR_MALFORMED, "slice %u band exceeds plane rows", idx); }
            const int32_t qp = (int32_t)fh->qp_base + (int32_t)h->qpd - 64;
            if (qp < 0 || qp > (int32_t)TC_QP_MAX) { return tc_fail(TC_ERR_MALFORMED, "slice %u effective qp %d", idx, qp); }
            if (fh->alpha_mode != 0u && p == 3u && h->k3 != 0u) { return tc_fail(TC_ERR_MALFORMED, "alpha slice %u k3 != 0", idx); }
            covered += h->h;
            idx++;
        }
    }
    if (idx != n) { return tc_fail(TC_ERR_MALFORMED, "%u trailing slices after full tiling", (unsigned)(n - idx)); }
    return TC_OK;
}

/* ------------------------------------------------------------ quantiser */

int tc_is_chroma(const tc_fhdr* fh, uint32_t p) { return fh->pixel_format != 3u && (p == 1u || p == 2u); }
int tc_is_alpha(const tc_fhdr* fh, uint32_t p) { return p == 3u && fh->alpha_mode != 0u; }

int32_t tc_qp_eff(const tc_fhdr* fh, const topos_frame_config* cfg, uint32_t p)
{
    int32_t qe = (int32_t)fh->qp_base + (tc_is_chroma(fh, p) ? cfg->qp_delta_chroma : cfg->qp_delta_luma)
               + (fh->bit_depth >= 12u ? 2 * ((int32_t)fh->bit_depth - 10) : 0);
    if (qe < 0) { qe = 0; }
    const int32_t ceil = fh->qp_base >= 64u ? 95 : 63;
    return qe > ceil ? ceil : qe;
}

static const uint32_t tc_tw[64] = TC_TW_INIT;

/* Q = max(1, (qm * qp_scale + 128) >> 8), half = Q >> 1, dead zone Q/4 on AC,
 * coefficient clamp 2^25 (scaled up for bit depth > 12); qwt = Q W 2^-32,
 * transposed, is exact in f64 (Q W < 2^40, and the scale is a power of two) */
void tc_qctx_build(tcg_qctx* q, const uint16_t* qm, uint32_t qp, uint32_t bd, uint32_t tbl)
{
    for (int i = 0; i < 64; ++i) {
        uint32_t Q = (uint32_t)(((uint64_t)qm[i] * tc_qp_scale_tab[tbl & 1u][qp] + 128u) >> 8);
        if (Q == 0u) { Q = 1u; }
        q->Q[i] = Q;
        q->qwt[(i & 7) * 8 + (i >> 3)] = (double)((uint64_t)Q * tc_tw[i]) * 0x1p-32;
        q->half[i] = Q >> 1;
        q->dz[i] = i == 0 ? 0u : Q / 4u;
        tc_qdiv32_consts(Q, &q->qmul[i], &q->qsh[i]);
    }
    q->bit_depth = bd;
    q->mid = 1 << (bd - 1u);
    q->maxv = (int32_t)((1u << bd) - 1u);
    q->fclamp = bd <= 12u ? (1 << 25)
                          : (int32_t)((((int64_t)1 << 25) * (int64_t)((1u << (bd - 1u)) - 1u)) / 2047);
    q->qlim_min = 0xFFFFFFFFu;
    for (int i = 0; i < 64; ++i) {
        q->qlim[i] = (uint32_t)q->fclamp / q->Q[i];
        if (q->qlim[i] < q->qlim_min) { q->qlim_min = q->qlim[i]; }
    }
}

/* alpha = clamp(2 * max(1, round(scale(qp) * 3 / 173056)), 3, 48)
 * (native/src/codec/deblock.c:71-80) */
uint32_t tc_deblock_alpha(uint32_t tbl, uint32_t qp)
{
    uint32_t step = (tc_qp_scale_tab[tbl & 1u][qp] * 3u + 173056u / 2u) / 173056u;
    if (step < 1u) { step = 1u; }
    uint32_t a = step * 2u;
    if (a < 3u) { a = 3u; }
    if (a > 48u) { a = 48u; }
    return a;
}

void tc_fill_output_info(const tc_fhdr* fh, topos_frame_output* info)
{
    info->visible_width = fh->visible_w;
    info->visible_height = fh->visible_h;
    info->coded_width = fh->coded_w;
    info->coded_height = fh->coded_h;
    info->plane_count = fh->plane_count;
    info->bit_depth = fh->bit_depth;
    info->profile = fh->profile;
    info->pixel_format = fh->pixel_format;
    info->alpha_mode = fh->alpha_mode;
    info->alpha_bit_depth = fh->alpha_bit_depth;
    info->color_range = fh->color_range;
    info->color_primaries = fh->color_primaries;
    info->color_transfer = fh->color_transfer;
    info->color_matrix = fh->color_matrix;
    info->chroma_siting = fh->chroma_siting;
    info->sar_num = fh->sar_num;
    info->sar_den = fh->sar_den;
    info->slice_count = fh->slice_count;
}

/* exported views of the tables above (include/topos_gpu.h) */
double tcx_ctx_log2(double x) { return tc_ctx_log2(x); }

void tcx_crc_tables(uint32_t crc256[256], uint32_t x2n32[32]) { tc_crc_tables(crc256, x2n32); }
