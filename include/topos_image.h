/* topos_image.h — the .toos (TPIM envelope) still-image read path of the
 * MI355X intra backend.
 *
 * Same names, struct layouts, status codes and check order as the reference
 * declarations they replace (reference file native/include/topos_image.h,
 * line numbers cited per item). The envelope is parsed on the host; the
 * PIXL packet is decoded by tc_frame_decode on the GPU, so decode results and
 * concealment are those of the frame path.
 */
#ifndef TOPOS_IMAGE_H
#define TOPOS_IMAGE_H

#include "topos_codec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* reference topos_image.h:25 */
#define TOPOS_IMAGE_ABI_VERSION 1

/* image-layer status codes (reference topos_image.h:28-35) */
#define TC_IMG_ERR_BAD_MAGIC (-100)
#define TC_IMG_ERR_BAD_PREAMBLE (-101)
#define TC_IMG_ERR_BAD_DIRECTORY (-102)
#define TC_IMG_ERR_UNKNOWN_CRITICAL (-103)
#define TC_IMG_ERR_CHUNK_CONFLICT (-104)
#define TC_IMG_ERR_METADATA_CONFLICT (-105)
#define TC_IMG_ERR_LIMIT (-106)
#define TC_IMG_ERR_IO_WRITE_FAILED (-107)

/* chunk FourCCs, big-endian (reference topos_image.h:38-56) */
#define TC_IMG_CHUNK_IDSC 0x49445343u
#define TC_IMG_CHUNK_PIXL 0x5049584Cu
#define TC_IMG_CHUNK_ICCP 0x49434350u
#define TC_IMG_CHUNK_OCIO 0x4F43494Fu
#define TC_IMG_CHUNK_XMP 0x584D5020u
#define TC_IMG_CHUNK_EXIF 0x45584946u
#define TC_IMG_CHUNK_THMB 0x54484D42u
#define TC_IMG_CHUNK_HASH 0x48415348u
#define TC_IMG_CHUNK_TMET 0x544D4554u
#define TC_IMG_CHUNK_TRWM 0x5452574Du

/* chunk flags (reference topos_image.h:59-62) */
#define TC_IMG_CHUNK_FLAG_CRITICAL 1u
#define TC_IMG_CHUNK_FLAG_OPTIONAL 2u
#define TC_IMG_CHUNK_FLAG_PRESERVE 4u
#define TC_IMG_CHUNK_FLAG_DROPPABLE 8u

/* fixed sizes and limits (reference topos_image.h:65-78) */
#define TC_IMG_PREAMBLE_SIZE 64u
#define TC_IMG_DIR_ENTRY_SIZE 32u
#define TC_IMG_IDSC_SIZE 128u
#define TC_IMG_MAX_CHUNKS 64u
#define TC_IMG_MAX_CHUNK_PIXL 268435456ull
#define TC_IMG_MAX_CHUNK_ICCP 8388608ull
#define TC_IMG_MAX_CHUNK_OCIO 4096ull
#define TC_IMG_MAX_CHUNK_XMP 16777216ull
#define TC_IMG_MAX_CHUNK_EXIF 16777216ull
#define TC_IMG_MAX_CHUNK_THMB 4194304ull
#define TC_IMG_MAX_CHUNK_HASH 36ull
#define TC_IMG_MAX_CHUNK_TRWM 256ull

/* IDSC field values (reference topos_image.h:81-104) */
#define TC_IMG_CHANNEL_MODEL_RGB 0u
#define TC_IMG_CHANNEL_MODEL_YUV 1u
#define TC_IMG_CHANNEL_MODEL_GRAY 2u
#define TC_IMG_CHANNEL_MODEL_CFA 3u
#define TC_IMG_SAMPLE_KIND_UINT 0u
#define TC_IMG_SAMPLE_KIND_HALF 1u
#define TC_IMG_LAYOUT_PLANAR 1u
#define TC_IMG_ALPHA_ABSENT 0u
#define TC_IMG_ALPHA_STRAIGHT 1u
#define TC_IMG_ALPHA_PREMULT 2u
#define TC_IMG_CODEC_ID_TPIC 0u
#define TC_IMG_PROFILE_PREVIEW 0u
#define TC_IMG_PROFILE_HQ 1u
#define TC_IMG_PROFILE_XQ 2u
#define TC_IMG_PROFILE_RAW 3u
#define TC_IMG_PROFILE_HALF_FLOAT 4u
#define TC_IMG_IDSC_FLAG_ORIENTATION_NORMALIZED 0x1u

/* reference topos_image.h:110-126 */
typedef struct topos_image_preamble {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t file_version_major;
    uint16_t file_version_minor;
    uint32_t flags;
    uint64_t file_size;
    uint64_t directory_offset;
    uint32_t directory_entry_size;
    uint32_t directory_count;
    uint32_t primary_image_index;
    uint32_t compatibility_flags;
    uint32_t header_crc32;
    uint32_t directory_crc32;
    uint32_t reserved[4];
} topos_image_preamble;

/* reference topos_image.h:128-135 */
typedef struct topos_image_dir_entry {
    uint32_t chunk_type;
    uint32_t chunk_flags;
    uint64_t chunk_offset;
    uint64_t chunk_size;
    uint32_t chunk_crc32;
    uint32_t reserved;
} topos_image_dir_entry;

/* reference topos_image.h:138-175 */
typedef struct topos_image_idsc {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t idsc_version_major;
    uint16_t idsc_version_minor;
    uint32_t flags;
    int32_t display_x_min, display_y_min, display_x_max, display_y_max;
    int32_t data_x_min, data_y_min, data_x_max, data_y_max;
    uint32_t orientation;
    uint32_t pixel_aspect_num;
    uint32_t pixel_aspect_den;
    uint8_t channel_model;
    uint8_t channel_count;
    uint8_t sample_kind;
    uint8_t valid_bit_depth;
    uint8_t container_bit_depth;
    uint8_t storage_layout;
    uint8_t subsampling;
    uint8_t chroma_siting;
    uint8_t alpha_presence;
    uint8_t alpha_mode;
    uint8_t alpha_bit_depth;
    uint8_t alpha_max_abs_err;
    uint8_t codec_id;
    uint8_t payload_major;
    uint8_t payload_minor;
    uint8_t image_profile;
    uint8_t codec_profile;
    uint8_t pixel_format;
    uint8_t color_primaries;
    uint8_t color_transfer;
    uint8_t color_matrix;
    uint8_t color_range;
    uint8_t iccp_ref;
    uint8_t ocio_ref;
    uint32_t payload_header_crc32;
} topos_image_idsc;

/* reference topos_image.h:178-198 */
typedef struct topos_image_info {
    uint32_t struct_size;
    uint32_t abi_version;
    topos_image_preamble preamble;
    topos_image_idsc idsc;
    uint64_t pixl_offset;
    uint64_t pixl_size;
    uint16_t visible_width;
    uint16_t visible_height;
    uint8_t plane_count;
    uint8_t has_alpha;
    uint8_t bit_depth;
    uint8_t profile;
    uint8_t pixel_format;
    uint8_t alpha_mode;
    uint8_t alpha_bit_depth;
    uint8_t alpha_premultiplied;
    uint32_t reserved[8];
} topos_image_info;

/* reference topos_image.h:203 */
const char* tc_image_status_message(int32_t status);
/* reference topos_image.h:208: preamble, directory, IDSC and the packet
 * header, cross-checked; no pixel allocation */
int32_t tc_image_probe(const topos_io* io, topos_image_info* out);
/* reference topos_image.h:212-213: probe + streamed CRC of every chunk;
 * TC_IMG_VALIDATE_DEEP adds a structure parse of the packet */
#define TC_IMG_VALIDATE_DEEP 1u
int32_t tc_image_validate(const topos_io* io, uint32_t flags, topos_image_info* out);
/* reference topos_image.h:219-221: the PIXL packet through tc_frame_decode */
int32_t tc_image_decode(const topos_io* io, const topos_plane_view planes[TC_FRAME_MAX_PLANES],
                        topos_frame_output* out_info);
/* reference topos_image.h:231-235: target-size decode of the PIXL packet
 * (the reduced request when the target is a fixed ratio of the source, else
 * the scaled one); out_info holds the target geometry */
int32_t tc_image_decode_preview(const topos_io* io, uint32_t target_width, uint32_t target_height,
                                const topos_plane_view planes[TC_FRAME_MAX_PLANES], topos_frame_output* out_info);
/* reference topos_image.h:239-240 */
int32_t tc_image_query_decode_buffer(const topos_image_info* info, uint32_t plane, uint32_t* width,
                                     uint32_t* height);

#ifdef __cplusplus
}
#endif
#endif
