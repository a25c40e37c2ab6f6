/* .toos (TPIM envelope) read path: preamble, chunk directory, IDSC and the
 * IDSC <-> packet header cross-check, in the reference's check order
 * (native/src/image/image_reader.c:63-297, image_container.c:31-345,647-801);
 * the PIXL packet is decoded by tc_frame_decode (image_reader.c:411-489).
 * Layouts are big-endian at fixed offsets (docs/image/topos_image_file_spec_v0.md
 * sections 3, 4 and 6).
 */
#include <stdlib.h>
#include <string.h>

#include "topos_image.h"
#include "tc_host.h"

#define TPIM_MAGIC 0x5450494Du

static uint32_t rd16(const uint8_t* p) { return ((uint32_t)p[0] << 8) | p[1]; }
static uint32_t rd32(const uint8_t* p)
{
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
static uint64_t rd64(const uint8_t* p) { return ((uint64_t)rd32(p) << 32) | rd32(p + 4); }

static int all_zero(const uint8_t* p, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        if (p[i] != 0u) { return 0; }
    }
    return 1;
}

static const char* fourcc(uint32_t t, char out[5])
{
    out[0] = (char)(t >> 24); out[1] = (char)(t >> 16); out[2] = (char)(t >> 8); out[3] = (char)t; out[4] = 0;
    return out;
}

const char* tc_image_status_message(int32_t status)
{
    switch (status) {
    case TC_IMG_ERR_BAD_MAGIC: return "bad magic (not TPIM)";
    case TC_IMG_ERR_BAD_PREAMBLE: return "bad preamble";
    case TC_IMG_ERR_BAD_DIRECTORY: return "bad chunk directory";
    case TC_IMG_ERR_UNKNOWN_CRITICAL: return "unknown critical chunk";
    case TC_IMG_ERR_CHUNK_CONFLICT: return "IDSC/PIXL conflict";
    case TC_IMG_ERR_METADATA_CONFLICT: return "color metadata conflict";
    case TC_IMG_ERR_LIMIT: return "image limit exceeded";
    case TC_IMG_ERR_IO_WRITE_FAILED: return "atomic write failed";
    case TC_OK: return "ok";
    case TC_WARN_CONCEALED: return "ok (concealed slices)";
    case TC_ERR_INVALID_ARGUMENT: return "invalid argument";
    case TC_ERR_OUT_OF_MEMORY: return "out of memory";
    case TC_ERR_UNSUPPORTED_VERSION: return "unsupported version";
    case TC_ERR_UNSUPPORTED_PROFILE: return "unsupported profile";
    case TC_ERR_UNSUPPORTED_PIXEL_FORMAT: return "unsupported pixel format";
    case TC_ERR_UNSUPPORTED_MATRIX: return "unsupported matrix";
    case TC_ERR_UNSUPPORTED_ALPHA_MODE: return "unsupported alpha mode";
    case TC_ERR_LIMIT_EXCEEDED: return "limit exceeded";
    case TC_ERR_MALFORMED: return "malformed data";
    case TC_ERR_TRUNCATED: return "truncated data";
    case TC_ERR_CHECKSUM_MISMATCH: return "checksum mismatch";
    case TC_ERR_STATE: return "invalid state";
    case TC_ERR_CANCELLED: return "cancelled";
    case TC_ERR_IO: return "io error";
    case TC_ERR_BUFFER_TOO_SMALL: return "buffer too small";
    case TC_ERR_NOT_IMPLEMENTED: return "not implemented";
    default: return "unknown status";
    }
}

static int32_t read_at(const topos_io* io, uint64_t off, void* buf, size_t n)
{
    if (io == NULL || io->read == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "image io: read callback missing"); }
    if (n == 0u) { return TC_OK; }
    const int32_t rc = io->read(io->ctx, off, buf, n);
    if (rc != TC_OK) {
        return tc_fail(TC_ERR_IO, "image read_at(%llu,%zu) failed: %d", (unsigned long long)off, n, (int)rc);
    }
    return TC_OK;
}

/* length limits of the known optional chunks (image_container.c:31-46) */
static int chunk_limit(uint32_t type, uint64_t* limit)
{
    switch (type) {
    case TC_IMG_CHUNK_ICCP: *limit = TC_IMG_MAX_CHUNK_ICCP; return 1;
    case TC_IMG_CHUNK_OCIO: *limit = TC_IMG_MAX_CHUNK_OCIO; return 1;
    case TC_IMG_CHUNK_XMP: *limit = TC_IMG_MAX_CHUNK_XMP; return 1;
    case TC_IMG_CHUNK_EXIF: *limit = TC_IMG_MAX_CHUNK_EXIF; return 1;
    case TC_IMG_CHUNK_THMB: *limit = TC_IMG_MAX_CHUNK_THMB; return 1;
    case TC_IMG_CHUNK_HASH: *limit = TC_IMG_MAX_CHUNK_HASH; return 1;
    case TC_IMG_CHUNK_TMET: *limit = 128u; return 1;
    case TC_IMG_CHUNK_TRWM: *limit = TC_IMG_MAX_CHUNK_TRWM; return 1;
    default: return 0;
    }
}

/* 64-byte preamble: magic, CRC over the block with its CRC field zeroed, then
 * the field rules (image_container.c:82-162) */
static int32_t preamble_parse(const uint8_t in[64], topos_image_preamble* p)
{
    memset(p, 0, sizeof(*p));
    p->struct_size = (uint32_t)sizeof(*p);
    p->abi_version = TOPOS_IMAGE_ABI_VERSION;
    if (rd32(in) != TPIM_MAGIC) { return tc_fail(TC_IMG_ERR_BAD_MAGIC, "preamble magic mismatch (not TPIM)"); }
    p->file_version_major = (uint16_t)rd16(in + 4);
    p->file_version_minor = (uint16_t)rd16(in + 6);
    p->flags = rd32(in + 12);
    p->file_size = rd64(in + 16);
    p->directory_offset = rd64(in + 24);
    p->directory_entry_size = rd32(in + 32);
    p->directory_count = rd32(in + 36);
    p->primary_image_index = rd32(in + 40);
    p->compatibility_flags = rd32(in + 44);
    p->header_crc32 = rd32(in + 48);
    p->directory_crc32 = rd32(in + 52);
    uint8_t z[64];
    memcpy(z, in, 64);
    memset(z + 48, 0, 4);
    if (tc_crc32(z, 64) != p->header_crc32) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "preamble header_crc32 mismatch"); }
    if (p->file_version_major != 1u) {
        return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "unsupported file_version_major=%u", (unsigned)p->file_version_major);
    }
    if (rd32(in + 8) != TC_IMG_PREAMBLE_SIZE) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "preamble_size != 64"); }
    if (p->flags != 0u) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "preamble flags undefined bits set"); }
    if (p->file_version_minor == 0u && (p->compatibility_flags & ~0x3u) != 0u) {
        return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "compatibility_flags undefined bits set (minor=0)");
    }
    if (p->directory_entry_size != TC_IMG_DIR_ENTRY_SIZE) {
        return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "directory_entry_size != 32");
    }
    if (p->directory_count < 2u || p->directory_count > TC_IMG_MAX_CHUNKS) {
        return tc_fail(TC_ERR_LIMIT_EXCEEDED, "directory_count=%u out of [2,%u]", (unsigned)p->directory_count,
                       (unsigned)TC_IMG_MAX_CHUNKS);
    }
    if (p->primary_image_index != 0u) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "primary_image_index != 0 (v1)"); }
    if (!all_zero(in + 56, 8)) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "preamble reserved0 != 0"); }
    return TC_OK;
}

/* 32-byte directory entry (image_container.c:180-214) */
static int32_t entry_parse(const uint8_t in[32], topos_image_dir_entry* e)
{
    memset(e, 0, sizeof(*e));
    e->chunk_type = rd32(in);
    e->chunk_flags = rd32(in + 4);
    e->chunk_offset = rd64(in + 8);
    e->chunk_size = rd64(in + 16);
    e->chunk_crc32 = rd32(in + 24);
    if (rd32(in + 28) != 0u) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "dir entry reserved != 0"); }
    if (e->chunk_size == 0u) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "dir entry chunk_size == 0"); }
    const uint32_t both = TC_IMG_CHUNK_FLAG_CRITICAL | TC_IMG_CHUNK_FLAG_OPTIONAL;
    if ((e->chunk_flags & both) == both) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "dir entry critical|optional both set");
    }
    if (e->chunk_flags & ~0xFu) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "dir entry undefined flag bits"); }
    return TC_OK;
}

/* bounds, no overlap with the directory, ascending and disjoint, no
 * duplicate critical chunk (image_reader.c:74-132) */
static int32_t directory_scan(const uint8_t* dir, uint32_t count, uint64_t file_size, uint64_t dir_off,
                              topos_image_dir_entry* ents)
{
    char cc[5];
    uint64_t prev_end = 0;
    const uint64_t dir_end = dir_off + (uint64_t)count * TC_IMG_DIR_ENTRY_SIZE;
    for (uint32_t i = 0; i < count; ++i) {
        int32_t rc = entry_parse(dir + (size_t)i * TC_IMG_DIR_ENTRY_SIZE, &ents[i]);
        if (rc != TC_OK) { return rc; }
        const topos_image_dir_entry* e = &ents[i];
        const uint64_t end = e->chunk_offset + e->chunk_size;
        if (e->chunk_offset < TC_IMG_PREAMBLE_SIZE || end < e->chunk_offset || end > file_size) {
            return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "chunk '%s' offset/size out of bounds", fourcc(e->chunk_type, cc));
        }
        if (e->chunk_offset < dir_end && end > dir_off) {
            return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "chunk overlaps directory");
        }
        if (e->chunk_offset < prev_end) {
            return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "directory not ascending / chunk overlap at entry %u", i);
        }
        prev_end = end;
    }
    for (uint32_t i = 0; i < count; ++i) {
        if (!(ents[i].chunk_flags & TC_IMG_CHUNK_FLAG_CRITICAL)) { continue; }
        for (uint32_t j = i + 1u; j < count; ++j) {
            if ((ents[j].chunk_flags & TC_IMG_CHUNK_FLAG_CRITICAL) && ents[j].chunk_type == ents[i].chunk_type) {
                return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "duplicate critical chunk '%s'", fourcc(ents[i].chunk_type, cc));
            }
        }
    }
    return TC_OK;
}

/* IDSC field rules without the packet (image_container.c:346-543) */
static int32_t idsc_rules(const topos_image_idsc* s)
{
    if (s->flags & ~TC_IMG_IDSC_FLAG_ORIENTATION_NORMALIZED) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC flags undefined bits");
    }
    if (s->display_x_max < s->display_x_min || s->display_y_max < s->display_y_min) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC display window inverted");
    }
    if (s->data_x_max < s->data_x_min || s->data_y_max < s->data_y_min) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC data window inverted");
    }
    if (s->display_x_min > s->data_x_min || s->display_y_min > s->data_y_min || s->display_x_max < s->data_x_max ||
        s->display_y_max < s->data_y_max) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "display window does not contain data window");
    }
    if (s->orientation < 1u || s->orientation > 8u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC orientation %u not in 1..8", (unsigned)s->orientation);
    }
    if (s->pixel_aspect_num < 1u || s->pixel_aspect_den < 1u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC pixel_aspect < 1:1");
    }
    if (s->channel_model > TC_IMG_CHANNEL_MODEL_CFA) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC channel_model=%u reserved", (unsigned)s->channel_model);
    }
    if (s->channel_count != 3u && s->channel_count != 4u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC channel_count=%u invalid", (unsigned)s->channel_count);
    }
    if (s->sample_kind != TC_IMG_SAMPLE_KIND_UINT && s->sample_kind != TC_IMG_SAMPLE_KIND_HALF) {
        return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "IDSC sample_kind=%u reserved", (unsigned)s->sample_kind);
    }
    const uint8_t vb = s->valid_bit_depth, cb = s->container_bit_depth;
    if (vb != 8u && vb != 10u && vb != 12u && vb != 16u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC valid_bit_depth=%u invalid", (unsigned)vb);
    }
    if (cb != 10u && cb != 12u && cb != 16u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC container_bit_depth=%u invalid", (unsigned)cb);
    }
    if (vb > cb) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC valid_bit_depth > container_bit_depth"); }
    if (s->storage_layout != TC_IMG_LAYOUT_PLANAR) {
        return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "IDSC storage_layout=%u reserved", (unsigned)s->storage_layout);
    }
    if (s->subsampling > 2u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC subsampling=%u reserved", (unsigned)s->subsampling);
    }
    if (s->alpha_presence > TC_IMG_ALPHA_PREMULT) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC alpha_presence=%u reserved", (unsigned)s->alpha_presence);
    }
    if (s->alpha_mode > 2u) {
        return tc_fail(TC_ERR_UNSUPPORTED_ALPHA_MODE, "IDSC alpha_mode=%u reserved", (unsigned)s->alpha_mode);
    }
    if (s->alpha_presence == TC_IMG_ALPHA_ABSENT) {
        if (s->alpha_mode != 0u || s->alpha_bit_depth != 0u || s->alpha_max_abs_err != 0u) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC alpha fields set while absent");
        }
        if (s->channel_count != 3u && s->channel_model != TC_IMG_CHANNEL_MODEL_CFA) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC channel_count=4 while alpha absent");
        }
        if (s->channel_model == TC_IMG_CHANNEL_MODEL_CFA && s->channel_count != 4u) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC CFA requires channel_count=4");
        }
    } else {
        if (s->alpha_mode == 0u) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC alpha present but alpha_mode=0"); }
        if (s->alpha_mode == 1u) {
            if (s->alpha_bit_depth != 16u || s->alpha_max_abs_err != 0u) {
                return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC alpha mode1 requires 16-bit/err=0");
            }
        } else if (s->alpha_bit_depth != 8u && s->alpha_bit_depth != 10u && s->alpha_bit_depth != 12u) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC alpha mode2 depth invalid");
        }
        if (s->channel_count != 4u) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC channel_count=3 while alpha present");
        }
    }
    if (s->codec_id != TC_IMG_CODEC_ID_TPIC) {
        return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "IDSC codec_id=%u reserved (v1 TPIC)", (unsigned)s->codec_id);
    }
    if (s->payload_major < 1u || s->payload_major > 8u) {
        return tc_fail(TC_ERR_UNSUPPORTED_VERSION, "IDSC payload_major=%u not in supported TPIC V1..V8",
                       (unsigned)s->payload_major);
    }
    if (s->image_profile > TC_IMG_PROFILE_HALF_FLOAT) {
        return tc_fail(TC_ERR_UNSUPPORTED_PROFILE, "IDSC image_profile=%u reserved in v1", (unsigned)s->image_profile);
    }
    if (s->sample_kind == TC_IMG_SAMPLE_KIND_HALF) {
        if (s->image_profile != TC_IMG_PROFILE_HALF_FLOAT) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC sample_kind=HALF requires image_profile 4");
        }
        if (vb != 16u || cb != 16u) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC HALF requires valid/container_bit_depth 16");
        }
        if (s->channel_model != TC_IMG_CHANNEL_MODEL_RGB) {
            return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC HALF requires GBR channel_model (RGB)");
        }
        if (s->subsampling != 0u) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC HALF requires 4:4:4"); }
        if (s->alpha_mode == 2u) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC HALF alpha mode2 undefined"); }
    } else if (s->image_profile == TC_IMG_PROFILE_HALF_FLOAT) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "IDSC image_profile 4 requires sample_kind=HALF");
    }
    if (s->iccp_ref > 1u || s->ocio_ref > 1u) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC iccp/ocio_ref invalid"); }
    return TC_OK;
}

/* 128-byte IDSC (image_container.c:268-344) */
static int32_t idsc_parse(const uint8_t in[128], topos_image_idsc* s)
{
    memset(s, 0, sizeof(*s));
    s->struct_size = (uint32_t)sizeof(*s);
    s->abi_version = TOPOS_IMAGE_ABI_VERSION;
    if (rd32(in) != TC_IMG_CHUNK_IDSC) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC magic mismatch"); }
    s->idsc_version_major = (uint16_t)rd16(in + 4);
    s->idsc_version_minor = (uint16_t)rd16(in + 6);
    s->flags = rd32(in + 12);
    s->display_x_min = (int32_t)rd32(in + 16);
    s->display_y_min = (int32_t)rd32(in + 20);
    s->display_x_max = (int32_t)rd32(in + 24);
    s->display_y_max = (int32_t)rd32(in + 28);
    s->data_x_min = (int32_t)rd32(in + 32);
    s->data_y_min = (int32_t)rd32(in + 36);
    s->data_x_max = (int32_t)rd32(in + 40);
    s->data_y_max = (int32_t)rd32(in + 44);
    s->orientation = rd32(in + 48);
    s->pixel_aspect_num = rd32(in + 52);
    s->pixel_aspect_den = rd32(in + 56);
    s->channel_model = in[60];
    s->channel_count = in[61];
    s->sample_kind = in[62];
    s->valid_bit_depth = in[63];
    s->container_bit_depth = in[64];
    s->storage_layout = in[65];
    s->subsampling = in[66];
    s->chroma_siting = in[67];
    s->alpha_presence = in[68];
    s->alpha_mode = in[69];
    s->alpha_bit_depth = in[70];
    s->alpha_max_abs_err = in[71];
    s->codec_id = in[72];
    s->payload_major = in[73];
    s->payload_minor = in[74];
    s->image_profile = in[75];
    s->codec_profile = in[76];
    s->pixel_format = in[77];
    s->color_primaries = in[78];
    s->color_transfer = in[79];
    s->color_matrix = in[80];
    s->color_range = in[81];
    s->iccp_ref = in[82];
    s->ocio_ref = in[83];
    s->payload_header_crc32 = rd32(in + 84);
    if (s->idsc_version_major != 1u || s->idsc_version_minor != 0u) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "unsupported IDSC version %u.%u", (unsigned)s->idsc_version_major,
                       (unsigned)s->idsc_version_minor);
    }
    if (rd32(in + 8) != TC_IMG_IDSC_SIZE) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC idsc_size != 128"); }
    if (!all_zero(in + 88, 40)) { return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC reserved bytes != 0"); }
    return idsc_rules(s);
}

/* IDSC <-> packet header (image_container.c:647-801) */
static int32_t idsc_crosscheck(const topos_image_idsc* s, const tc_fhdr* fh, const uint8_t raw[TC_HDR_SIZE])
{
    const int64_t dw = (int64_t)s->data_x_max - s->data_x_min + 1, dh = (int64_t)s->data_y_max - s->data_y_min + 1;
    if (dw != (int64_t)fh->visible_w || dh != (int64_t)fh->visible_h) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "data window %lldx%lld != packet visible %ux%u", (long long)dw,
                       (long long)dh, (unsigned)fh->visible_w, (unsigned)fh->visible_h);
    }
    if (s->display_x_min > s->data_x_min || s->display_y_min > s->data_y_min || s->display_x_max < s->data_x_max ||
        s->display_y_max < s->data_y_max) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "display window does not contain data window");
    }
    if (fh->pixel_format != s->pixel_format) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "pixel_format != IDSC"); }
    if (fh->bit_depth != s->valid_bit_depth) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "bit_depth != IDSC valid_bit_depth");
    }
    if (fh->profile != s->codec_profile) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "codec profile != IDSC"); }
    int ok = 0;
    switch (s->image_profile) {
    case TC_IMG_PROFILE_PREVIEW:
        ok = fh->pixel_format == 0u && fh->profile == 3u && (fh->bit_depth == 10u || fh->bit_depth == 12u);
        break;
    case TC_IMG_PROFILE_HQ: ok = fh->pixel_format == 2u && fh->bit_depth == 10u && fh->profile == 5u; break;
    case TC_IMG_PROFILE_XQ:
        ok = fh->pixel_format == 2u && fh->bit_depth == 12u && (fh->profile == 5u || fh->profile == 6u);
        break;
    case TC_IMG_PROFILE_RAW:
        ok = fh->pixel_format == 3u && (fh->bit_depth == 12u || fh->bit_depth == 16u) && fh->profile == 7u;
        break;
    case TC_IMG_PROFILE_HALF_FLOAT:
        ok = fh->pixel_format == 2u && fh->bit_depth == 16u && fh->profile == 5u &&
             s->sample_kind == TC_IMG_SAMPLE_KIND_HALF;
        break;
    default: ok = 0; break;
    }
    if (!ok) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "image_profile %u != codec combo (pf=%u bit_depth=%u profile=%u)",
                       (unsigned)s->image_profile, (unsigned)fh->pixel_format, (unsigned)fh->bit_depth,
                       (unsigned)fh->profile);
    }
    if (fh->alpha_mode != s->alpha_mode || fh->alpha_bit_depth != s->alpha_bit_depth) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "alpha mode/depth != IDSC");
    }
    const uint32_t premul = (fh->flags & 1u) ? TC_IMG_ALPHA_PREMULT : TC_IMG_ALPHA_STRAIGHT;
    if (s->alpha_presence != TC_IMG_ALPHA_ABSENT && s->alpha_presence != premul) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "alpha association != IDSC");
    }
    if (fh->color_range != s->color_range || fh->color_primaries != s->color_primaries ||
        fh->color_transfer != s->color_transfer || fh->color_matrix != s->color_matrix ||
        fh->chroma_siting != s->chroma_siting) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "color tags != IDSC");
    }
    const uint32_t square = fh->sar_num == 0u && fh->sar_den == 0u;
    const uint32_t wn = square ? 1u : fh->sar_num, wd = square ? 1u : fh->sar_den;
    if (wn != s->pixel_aspect_num || wd != s->pixel_aspect_den) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "sar != IDSC pixel_aspect");
    }
    if (rd32(raw + 49) != s->payload_header_crc32) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "payload_header_crc32 != IDSC binding digest");
    }
    if (fh->major != s->payload_major || fh->minor != s->payload_minor) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "payload version != IDSC");
    }
    const uint8_t model = fh->pixel_format == 2u ? TC_IMG_CHANNEL_MODEL_RGB
                        : fh->pixel_format == 3u ? TC_IMG_CHANNEL_MODEL_CFA : TC_IMG_CHANNEL_MODEL_YUV;
    if (s->channel_model != model) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "channel_model != pixel_format"); }
    if (s->channel_count != fh->plane_count) { return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "channel_count != plane_count"); }
    if ((fh->alpha_mode != 0u) != (s->alpha_presence != TC_IMG_ALPHA_ABSENT)) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "alpha presence mismatch");
    }
    return TC_OK;
}

/* probe (image_reader.c:135-297) */
static int32_t probe(const topos_io* io, topos_image_info* out)
{
    if (out == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "image probe: null out"); }
    memset(out, 0, sizeof(*out));
    out->struct_size = (uint32_t)sizeof(*out);
    out->abi_version = TOPOS_IMAGE_ABI_VERSION;
    if (io == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "image probe: null io"); }
    if (io->length < TC_IMG_PREAMBLE_SIZE) {
        return tc_fail(TC_ERR_TRUNCATED, "file shorter than preamble (%llu)", (unsigned long long)io->length);
    }
    uint8_t pre[TC_IMG_PREAMBLE_SIZE];
    int32_t rc = read_at(io, 0, pre, sizeof(pre));
    if (rc != TC_OK) { return rc; }
    if ((rc = preamble_parse(pre, &out->preamble)) != TC_OK) { return rc; }
    const topos_image_preamble* P = &out->preamble;
    if (P->file_size != io->length) { return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "preamble file_size != actual"); }
    const uint64_t dir_len = (uint64_t)P->directory_count * TC_IMG_DIR_ENTRY_SIZE;
    const uint64_t dir_end = P->directory_offset + dir_len;
    if (P->directory_offset < TC_IMG_PREAMBLE_SIZE || dir_end < P->directory_offset || dir_end > P->file_size) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "directory out of bounds");
    }
    uint8_t dir[TC_IMG_MAX_CHUNKS * TC_IMG_DIR_ENTRY_SIZE];
    if ((rc = read_at(io, P->directory_offset, dir, (size_t)dir_len)) != TC_OK) { return rc; }
    if (tc_crc32(dir, (size_t)dir_len) != P->directory_crc32) {
        return tc_fail(TC_IMG_ERR_BAD_PREAMBLE, "directory_crc32 mismatch");
    }
    topos_image_dir_entry ents[TC_IMG_MAX_CHUNKS];
    if ((rc = directory_scan(dir, P->directory_count, P->file_size, P->directory_offset, ents)) != TC_OK) { return rc; }
    uint32_t idsc_i = TC_IMG_MAX_CHUNKS, pixl_i = TC_IMG_MAX_CHUNKS;
    int has_iccp = 0, has_ocio = 0;
    char cc[5];
    for (uint32_t i = 0; i < P->directory_count; ++i) {
        const topos_image_dir_entry* e = &ents[i];
        if ((e->chunk_flags & TC_IMG_CHUNK_FLAG_CRITICAL) && e->chunk_type != TC_IMG_CHUNK_IDSC &&
            e->chunk_type != TC_IMG_CHUNK_PIXL) {
            return tc_fail(TC_IMG_ERR_UNKNOWN_CRITICAL, "unknown critical chunk '%s'", fourcc(e->chunk_type, cc));
        }
        uint64_t limit = 0;
        if (chunk_limit(e->chunk_type, &limit) && e->chunk_size > limit) {
            return tc_fail(TC_IMG_ERR_LIMIT, "chunk '%s' size %llu > limit %llu", fourcc(e->chunk_type, cc),
                           (unsigned long long)e->chunk_size, (unsigned long long)limit);
        }
        if (e->chunk_type == TC_IMG_CHUNK_IDSC) { idsc_i = i; }
        else if (e->chunk_type == TC_IMG_CHUNK_PIXL) { pixl_i = i; }
        else if (e->chunk_type == TC_IMG_CHUNK_ICCP) { has_iccp = 1; }
        else if (e->chunk_type == TC_IMG_CHUNK_OCIO) { has_ocio = 1; }
    }
    if (idsc_i == TC_IMG_MAX_CHUNKS || pixl_i == TC_IMG_MAX_CHUNKS) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "missing required chunk (IDSC/PIXL)");
    }
    if (ents[idsc_i].chunk_size != TC_IMG_IDSC_SIZE) {
        return tc_fail(TC_IMG_ERR_BAD_DIRECTORY, "IDSC chunk size %llu != 128",
                       (unsigned long long)ents[idsc_i].chunk_size);
    }
    uint8_t idsc[TC_IMG_IDSC_SIZE];
    if ((rc = read_at(io, ents[idsc_i].chunk_offset, idsc, sizeof(idsc))) != TC_OK) { return rc; }
    if ((rc = idsc_parse(idsc, &out->idsc)) != TC_OK) { return rc; }
    if ((out->idsc.iccp_ref != 0u) != has_iccp) {
        return tc_fail(TC_IMG_ERR_METADATA_CONFLICT, "iccp_ref=%u but ICCP chunk %s", (unsigned)out->idsc.iccp_ref,
                       has_iccp ? "present" : "missing");
    }
    if ((out->idsc.ocio_ref != 0u) != has_ocio) {
        return tc_fail(TC_IMG_ERR_METADATA_CONFLICT, "ocio_ref=%u but OCIO chunk %s", (unsigned)out->idsc.ocio_ref,
                       has_ocio ? "present" : "missing");
    }
    out->pixl_offset = ents[pixl_i].chunk_offset;
    out->pixl_size = ents[pixl_i].chunk_size;
    if (out->pixl_size > TC_IMG_MAX_CHUNK_PIXL) { return tc_fail(TC_IMG_ERR_LIMIT, "PIXL size > 256 MiB"); }
    if (out->pixl_size < TC_HDR_SIZE) { return tc_fail(TC_ERR_TRUNCATED, "PIXL shorter than frame header"); }
    uint8_t raw[TC_HDR_SIZE];
    if ((rc = read_at(io, out->pixl_offset, raw, sizeof(raw))) != TC_OK) { return rc; }
    tc_fhdr fh;
    if ((rc = tc_fhdr_parse(raw, sizeof(raw), &fh)) != TC_OK) { return rc; }
    if ((rc = idsc_crosscheck(&out->idsc, &fh, raw)) != TC_OK) { return rc; }
    if ((uint64_t)fh.packet_size != out->pixl_size) {
        return tc_fail(TC_IMG_ERR_CHUNK_CONFLICT, "frame_packet_size %u != PIXL chunk size %llu",
                       (unsigned)fh.packet_size, (unsigned long long)out->pixl_size);
    }
    out->visible_width = fh.visible_w;
    out->visible_height = fh.visible_h;
    out->plane_count = fh.plane_count;
    out->has_alpha = fh.alpha_mode != 0u ? 1u : 0u;
    out->bit_depth = fh.bit_depth;
    out->profile = fh.profile;
    out->pixel_format = fh.pixel_format;
    out->alpha_mode = fh.alpha_mode;
    out->alpha_bit_depth = fh.alpha_bit_depth;
    out->alpha_premultiplied = (fh.flags & 1u) ? 1u : 0u;
    return TC_OK;
}

int32_t tc_image_probe(const topos_io* io, topos_image_info* out) { return probe(io, out); }

/* probe + streamed CRC of every chunk; DEEP adds the packet structure parse
 * (image_reader.c:304-380) */
int32_t tc_image_validate(const topos_io* io, uint32_t flags, topos_image_info* out)
{
    int32_t rc = probe(io, out);
    if (rc != TC_OK) { return rc; }
    const uint32_t count = out->preamble.directory_count;
    uint8_t dir[TC_IMG_MAX_CHUNKS * TC_IMG_DIR_ENTRY_SIZE];
    if ((rc = read_at(io, out->preamble.directory_offset, dir, (size_t)count * TC_IMG_DIR_ENTRY_SIZE)) != TC_OK) {
        return rc;
    }
    enum { BUF = 64 * 1024 };
    uint8_t* buf = (uint8_t*)malloc(BUF);
    if (buf == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "validate stream buffer alloc"); }
    char cc[5];
    for (uint32_t i = 0; i < count && rc == TC_OK; ++i) {
        topos_image_dir_entry e;
        if ((rc = entry_parse(dir + (size_t)i * TC_IMG_DIR_ENTRY_SIZE, &e)) != TC_OK) { break; }
        uint32_t crc = 0;
        for (uint64_t off = e.chunk_offset, left = e.chunk_size; left > 0u && rc == TC_OK;) {
            const size_t n = (size_t)(left > BUF ? BUF : left);
            if ((rc = read_at(io, off, buf, n)) == TC_OK) {
                crc = tc_crc32_update(crc, buf, n);
                off += n;
                left -= n;
            }
        }
        if (rc == TC_OK && crc != e.chunk_crc32) {
            rc = tc_fail(TC_ERR_CHECKSUM_MISMATCH, "chunk '%s' crc mismatch (store %08x calc %08x)",
                         fourcc(e.chunk_type, cc), (unsigned)e.chunk_crc32, (unsigned)crc);
        }
    }
    free(buf);
    if (rc != TC_OK) { return rc; }
    if (flags & TC_IMG_VALIDATE_DEEP) {
        uint8_t* pixl = (uint8_t*)malloc((size_t)out->pixl_size);
        if (pixl == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "validate deep PIXL alloc"); }
        rc = read_at(io, out->pixl_offset, pixl, (size_t)out->pixl_size);
        if (rc == TC_OK) {
            topos_frame_output info;
            memset(&info, 0, sizeof(info));
            rc = tc_frame_decode(pixl, (size_t)out->pixl_size, NULL, NULL, &info);
        }
        free(pixl);
    }
    return rc;
}

/* reference image_reader.c:382-401 */
int32_t tc_image_query_decode_buffer(const topos_image_info* info, uint32_t plane, uint32_t* width, uint32_t* height)
{
    if (info == NULL || info->struct_size != (uint32_t)sizeof(topos_image_info) ||
        info->abi_version != TOPOS_IMAGE_ABI_VERSION) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "query_decode_buffer: bad info");
    }
    topos_frame_output fo;
    memset(&fo, 0, sizeof(fo));
    fo.struct_size = (uint32_t)sizeof(fo);
    fo.abi_version = TOPOS_CODEC_ABI_VERSION;
    fo.visible_width = info->visible_width;
    fo.visible_height = info->visible_height;
    fo.plane_count = info->plane_count;
    fo.alpha_mode = info->alpha_mode;
    fo.bit_depth = info->bit_depth;
    fo.pixel_format = info->pixel_format;
    return tc_frame_plane_geometry(&fo, plane, width, height);
}

/* plane views: a zero struct_size is tolerated, otherwise size and codec ABI
 * version must match (image_reader.c:411-446) */
static int32_t views_check(const topos_plane_view planes[TC_FRAME_MAX_PLANES], uint16_t* ptrs[TC_FRAME_MAX_PLANES],
                           size_t strides[TC_FRAME_MAX_PLANES])
{
    if (planes == NULL) { return tc_fail(TC_ERR_INVALID_ARGUMENT, "image decode: null planes"); }
    for (uint32_t i = 0; i < TC_FRAME_MAX_PLANES; ++i) {
        ptrs[i] = NULL;
        strides[i] = 0;
        if (planes[i].struct_size != 0u) {
            if (planes[i].struct_size != (uint32_t)sizeof(topos_plane_view)) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT, "image decode: plane %u struct_size", i);
            }
            if (planes[i].abi_version != TOPOS_CODEC_ABI_VERSION) {
                return tc_fail(TC_ERR_INVALID_ARGUMENT, "image decode: plane %u abi_version %u != codec %u", i,
                               (unsigned)planes[i].abi_version, (unsigned)TOPOS_CODEC_ABI_VERSION);
            }
        }
        if (planes[i].pixels == NULL) { continue; }
        ptrs[i] = planes[i].pixels;
        strides[i] = planes[i].stride;
    }
    return TC_OK;
}

int32_t tc_image_decode(const topos_io* io, const topos_plane_view planes[TC_FRAME_MAX_PLANES],
                        topos_frame_output* out_info)
{
    topos_image_info info;
    int32_t rc = probe(io, &info);
    if (rc != TC_OK) { return rc; }
    uint16_t* ptrs[TC_FRAME_MAX_PLANES];
    size_t strides[TC_FRAME_MAX_PLANES];
    if ((rc = views_check(planes, ptrs, strides)) != TC_OK) { return rc; }
    uint8_t* pixl = (uint8_t*)malloc((size_t)info.pixl_size);
    if (pixl == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "image decode PIXL alloc"); }
    rc = read_at(io, info.pixl_offset, pixl, (size_t)info.pixl_size);
    if (rc == TC_OK) {
        if (out_info != NULL) {
            memset(out_info, 0, sizeof(*out_info));
            out_info->struct_size = (uint32_t)sizeof(*out_info);
            out_info->abi_version = TOPOS_CODEC_ABI_VERSION;
        }
        rc = tc_frame_decode(pixl, (size_t)info.pixl_size, (uint16_t* const*)ptrs, strides, out_info);
    }
    free(pixl);
    return rc;
}

/* reference native/src/image/image_reader.c:490-583: a target matching a
 * fixed ratio exactly goes through the reduced request, any other box through
 * the scaled one; plane views are checked against the target geometry */
int32_t tc_image_decode_preview(const topos_io* io, uint32_t target_width, uint32_t target_height,
                                const topos_plane_view planes[TC_FRAME_MAX_PLANES], topos_frame_output* out_info)
{
    if (planes == NULL || out_info == NULL || target_width == 0u || target_height == 0u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "image preview: invalid arguments");
    }
    topos_image_info info;
    int32_t rc = probe(io, &info);
    if (rc != TC_OK) { return rc; }
    if (target_width > (uint32_t)info.visible_width || target_height > (uint32_t)info.visible_height ||
        target_width > 65535u || target_height > 65535u) {
        return tc_fail(TC_ERR_INVALID_ARGUMENT, "image preview: target %ux%u exceeds source %ux%u", target_width,
                       target_height, (unsigned)info.visible_width, (unsigned)info.visible_height);
    }
    uint8_t* pixl = (uint8_t*)malloc((size_t)info.pixl_size);
    if (pixl == NULL) { return tc_fail(TC_ERR_OUT_OF_MEMORY, "image preview PIXL alloc"); }
    rc = read_at(io, info.pixl_offset, pixl, (size_t)info.pixl_size);
    topos_decode_request req;
    memset(&req, 0, sizeof(req));
    req.struct_size = (uint32_t)sizeof(req);
    req.abi_version = TOPOS_CODEC_ABI_VERSION;
    req.mode = TC_DECODE_MODE_SCALED;
    req.target_width = target_width;
    req.target_height = target_height;
    req.memory_type = TC_DECODE_MEMORY_CPU;
    for (uint32_t s = TC_DECODE_SCALE_HALF; s <= TC_DECODE_SCALE_EIGHTH; ++s) {
        uint32_t tw = 0, th = 0;
        if (tc_decode_scale_dimensions(info.visible_width, info.visible_height, (tc_decode_scale)s, &tw, &th) ==
                TC_OK && tw == target_width && th == target_height) {
            req.mode = TC_DECODE_MODE_REDUCED;
            req.scale = s;
            req.target_width = 0u;
            req.target_height = 0u;
            break;
        }
    }
    topos_frame_output tgt;
    memset(&tgt, 0, sizeof(tgt));
    if (rc == TC_OK) { rc = tc_frame_decode_request(pixl, (size_t)info.pixl_size, &req, NULL, NULL, &tgt); }
    if (rc == TC_OK && (tgt.plane_count == 0u || tgt.plane_count > TC_FRAME_MAX_PLANES)) {
        rc = tc_fail(TC_ERR_MALFORMED, "image preview: invalid source plane count");
    }
    uint16_t* ptrs[TC_FRAME_MAX_PLANES] = { NULL, NULL, NULL, NULL };
    size_t strides[TC_FRAME_MAX_PLANES] = { 0, 0, 0, 0 };
    for (uint32_t p = 0; rc == TC_OK && p < tgt.plane_count; ++p) {
        uint32_t w = 0, h = 0;
        rc = tc_frame_plane_geometry(&tgt, p, &w, &h);
        if (rc == TC_OK && (planes[p].struct_size != (uint32_t)sizeof(topos_plane_view) ||
                            planes[p].abi_version != TOPOS_CODEC_ABI_VERSION || planes[p].pixels == NULL)) {
            rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "image preview: plane %u view is invalid", p);
        }
        if (rc == TC_OK && planes[p].stride != 0u && planes[p].stride < (size_t)w) {
            rc = tc_fail(TC_ERR_INVALID_ARGUMENT, "image preview: plane %u stride %zu < width %u", p, planes[p].stride,
                         w);
        }
        ptrs[p] = planes[p].pixels;
        strides[p] = planes[p].stride;
    }
    if (rc == TC_OK) {
        rc = tc_frame_decode_request(pixl, (size_t)info.pixl_size, &req, ptrs, strides, &tgt);
        if (rc >= 0) { *out_info = tgt; }
    }
    free(pixl);
    return rc;
}
