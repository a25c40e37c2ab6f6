/* topos_codec.h — frame encode/decode C ABI of the MI355X intra backend.
 *
 * Drop-in for the frame section of the Topos C ABI: the structs below have the
 * same layout (ABI version 2) and the functions the same names, argument
 * meaning, status codes and error behaviour as the reference declarations they
 * replace (reference file native/include/topos_codec.h, line numbers cited per
 * item). Every call runs on the first visible AMD GPU; the calls are
 * thread-safe (device work is serialised per process) and a decoder context is
 * not re-entrant.
 */
#ifndef TOPOS_CODEC_H
#define TOPOS_CODEC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ABI generation of the public structs (reference topos_codec.h:31) */
#define TOPOS_CODEC_ABI_VERSION 2

/* status codes, frozen values (reference topos_codec.h:33-56) */
enum {
    TC_OK = 0,
    TC_WARN_CONCEALED = 1,
    TC_ERR_INVALID_ARGUMENT = -1,
    TC_ERR_OUT_OF_MEMORY = -2,
    TC_ERR_UNSUPPORTED_VERSION = -3,
    TC_ERR_UNSUPPORTED_PROFILE = -4,
    TC_ERR_UNSUPPORTED_PIXEL_FORMAT = -5,
    TC_ERR_UNSUPPORTED_MATRIX = -6,
    TC_ERR_UNSUPPORTED_ALPHA_MODE = -7,
    TC_ERR_LIMIT_EXCEEDED = -8,
    TC_ERR_MALFORMED = -9,
    TC_ERR_TRUNCATED = -10,
    TC_ERR_CHECKSUM_MISMATCH = -11,
    TC_ERR_STATE = -12,
    TC_ERR_CANCELLED = -13,
    TC_ERR_IO = -14,
    TC_ERR_BUFFER_TOO_SMALL = -15,
    TC_ERR_NOT_IMPLEMENTED = -16,
    TC_ERR_REFERENCE_INVALID = -17
};

#define TC_FRAME_MAX_PLANES 4
#define TC_FRAME_DEFAULT_SLICE_ROWS 16u
#define TC_FRAME_SLICE_OK 0u
#define TC_FRAME_SLICE_CONCEALED 1u
#define TC_MAX_SLICES 512

/* encoder configuration; all-zero (plus struct_size/abi_version) is a valid
 * minimal config (reference topos_codec.h:175-224). reserved[0] low byte
 * selects the entropy mode (8 = rANS2 intra product mode), reserved[5] the
 * output features of mode 8 (bit0 deblock, bit1 refined qp table). */
typedef struct topos_frame_config {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t visible_width;
    uint16_t visible_height;
    uint8_t profile;
    uint8_t pixel_format;      /* 0 YUV 4:2:2, 1 YUV 4:4:4, 2 GBR, 3 CFA (4 phase planes) */
    uint8_t bit_depth;
    uint8_t qmatrix_id;
    uint8_t qp_base;
    int8_t qp_delta_luma;
    int8_t qp_delta_chroma;
    uint8_t slice_rows;        /* block rows per slice; 0 = 16 */
    uint8_t alpha_mode;        /* 0 none, 1 lossless, 2 limited precision */
    uint8_t alpha_bit_depth;
    uint8_t alpha_premultiplied;
    uint8_t color_range;
    uint8_t color_primaries;
    uint8_t color_transfer;
    uint8_t color_matrix;
    uint8_t chroma_siting;
    uint16_t sar_num;
    uint16_t sar_den;
    uint32_t reserved[8];
} topos_frame_config;

/* planar uint16 input; strides in uint16 elements, 0 = tight
 * (reference topos_codec.h:226-237) */
typedef struct topos_frame_input {
    uint32_t struct_size;
    uint32_t abi_version;
    const uint16_t* planes[TC_FRAME_MAX_PLANES];
    size_t strides[TC_FRAME_MAX_PLANES];
    uint32_t reserved[8];
} topos_frame_input;

/* encode statistics (reference topos_codec.h:239-254) */
typedef struct topos_frame_stats {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t packet_size;
    uint32_t color_payload_bytes;
    uint32_t alpha_payload_bytes;
    uint32_t color_header_bytes;
    uint32_t alpha_header_bytes;
    uint16_t slice_count;
    uint8_t qp_base;
    uint8_t reserved8;
    uint16_t alpha_max_abs_error;
    uint16_t reserved16;
    uint32_t reserved[8];
} topos_frame_stats;

/* decode description + per-slice status (reference topos_codec.h:256-285) */
typedef struct topos_frame_output {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t visible_width;
    uint16_t visible_height;
    uint16_t coded_width;
    uint16_t coded_height;
    uint8_t plane_count;
    uint8_t bit_depth;
    uint8_t profile;
    uint8_t pixel_format;
    uint8_t alpha_mode;
    uint8_t alpha_bit_depth;
    uint8_t color_range;
    uint8_t color_primaries;
    uint8_t color_transfer;
    uint8_t color_matrix;
    uint8_t chroma_siting;
    uint16_t concealed_slices;
    uint16_t slice_count;
    uint16_t sar_num;
    uint16_t sar_den;
    uint8_t slice_status[TC_MAX_SLICES];
    uint32_t reserved[8];
} topos_frame_output;

/* batch packet reference (reference topos_codec.h:387-390) */
typedef struct topos_batch_packet {
    const uint8_t* data;
    size_t size;
} topos_batch_packet;

/* output plane view for decoder contexts (reference topos_codec.h:427-432) */
typedef struct topos_plane_view {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t* pixels;
    size_t stride;
} topos_plane_view;

/* caller byte source / sink (reference topos_codec.h:646-669): read exactly
 * len bytes at an absolute offset; write appends; seek_write overwrites len
 * bytes at an absolute offset (the MOV writer patches the mdat size with it).
 * Each returns TC_OK, or TC_ERR_IO on a short or failed transfer. */
typedef int32_t (*tc_io_read_fn)(void* ctx, uint64_t offset, void* buf, size_t len);
typedef int32_t (*tc_io_write_fn)(void* ctx, const void* data, size_t len);
typedef int32_t (*tc_io_seek_write_fn)(void* ctx, uint64_t offset, const void* data, size_t len);

typedef struct topos_io {
    uint32_t struct_size;
    uint32_t abi_version;
    void* ctx;
    tc_io_read_fn read;
    tc_io_write_fn write;
    tc_io_seek_write_fn seek_write;
    uint64_t length;
    uint32_t reserved[4];
} topos_io;

/* decoder context config (reference topos_codec.h:502-510) */
typedef struct topos_decoder_config {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t max_slice_workers;
    uint32_t reserved[7];
} topos_decoder_config;

typedef struct tc_decoder tc_decoder;

/* library version (reference topos_codec.h:58-67); the strings are static */
typedef struct topos_version_info {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t version_major;
    uint32_t version_minor;
    uint32_t version_patch;
    const char* git_commit;
    const char* build_target;
    uint32_t reserved[4];
} topos_version_info;

/* host CPU feature bits (reference topos_codec.h:69-80); informational only */
#define TOPOS_CPU_X86_AVX2 (1u << 0)
#define TOPOS_CPU_X86_AVX512F (1u << 1)
#define TOPOS_CPU_ARM_NEON (1u << 2)
#define TOPOS_CPU_X86_FMA (1u << 3)

typedef struct topos_cpu_features {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t flags;
    uint32_t reserved[5];
} topos_cpu_features;

/* codec capability bits (reference topos_codec.h:82-104) */
#define TC_CODEC_CAP_FULL_DECODE (1u << 0)
#define TC_CODEC_CAP_FIXED_REDUCED_DECODE (1u << 1)
#define TC_CODEC_CAP_AUTO_2K_DECODE (1u << 2)
#define TC_CODEC_CAP_CPU_SURFACE_OUTPUT (1u << 3)
#define TC_CODEC_CAP_HOST_VISIBLE_GPU_SURFACE (1u << 4)
#define TC_CODEC_CAP_V7A_BAND_DECODE (1u << 5)
#define TC_CODEC_CAP_V7B_SCALABLE_BASE (1u << 6)
#define TC_CODEC_CAP_DECODE_DROP_ALPHA (1u << 7)

#define TC_CODEC_MAX_DIM 16384u
#define TC_CODEC_AUTO_2K_MAX_DIM 2048u

typedef struct topos_codec_capabilities {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t flags;
    uint32_t max_width;
    uint32_t max_height;
    uint32_t auto2k_max_dim;
    uint32_t reserved[4];
} topos_codec_capabilities;

/* release state of the scalable (V7-B) layout (reference topos_codec.h:110-115) */
typedef enum topos_scalable_status {
    TC_SCALABLE_STATUS_EXPERIMENTAL = 0,
    TC_SCALABLE_STATUS_REFERENCE = 1,
    TC_SCALABLE_STATUS_ELIGIBLE = 2,
    TC_SCALABLE_STATUS_DEFAULT = 3
} topos_scalable_status;

/* reference topos_codec.h:125 */
int32_t tc_abi_version(void);
/* reference topos_codec.h:128: TC_ERR_INVALID_ARGUMENT when out == NULL */
int32_t tc_version(topos_version_info* out);
/* reference topos_codec.h:131 */
int32_t tc_query_cpu_features(topos_cpu_features* out);
/* reference topos_codec.h:135: the decode paths this library provides */
int32_t tc_query_capabilities(topos_codec_capabilities* out);
/* reference topos_codec.h:142-143: TC_OK when a stream of this profile /
 * pixel format / bit depth / alpha mode decodes here, else the
 * TC_ERR_UNSUPPORTED_* code of the first field that fails */
int32_t tc_query_support(uint32_t profile, uint32_t pixel_format, uint32_t bit_depth, uint32_t alpha_mode);
/* reference topos_codec.h:118, 122: V7-B is not written or read by this
 * library, so the state reported is TC_SCALABLE_STATUS_EXPERIMENTAL */
int32_t tc_query_scalable_status(topos_scalable_status* out);
const char* tc_scalable_status_name(topos_scalable_status status);
/* reference topos_codec.h:146 */
const char* tc_status_message(int32_t status);
/* reference topos_codec.h:150 */
const char* tc_last_error(void);

/* reference topos_codec.h:288 */
int32_t tc_frame_config_validate(const topos_frame_config* cfg);
/* reference topos_codec.h:292 */
size_t tc_frame_packet_bound(const topos_frame_config* cfg);
/* reference topos_codec.h:296-297: out untouched unless TC_OK */
int32_t tc_frame_encode(const topos_frame_config* cfg, const topos_frame_input* input,
                        uint8_t* out, size_t out_cap, topos_frame_stats* stats);
/* reference topos_codec.h:315-328: target-size encode. Deterministic qp search
 * in [qp_min, qp_max] (strict by default: the smallest qp whose packet fits
 * target_bytes; legacy linear when a probe sequence breaks monotonicity or
 * the 24-probe budget runs out; TOPOS_SIZED_SEARCH=linear|r6|strict selects
 * the search once per process), seeded by the process's last two working
 * points; the packet equals tc_frame_encode at *qp_used. out untouched unless
 * TC_OK; TC_ERR_BUFFER_TOO_SMALL reports the final size in stats. RAW
 * (pixel format 3) and entropy modes 9-11 return TC_ERR_NOT_IMPLEMENTED. */
int32_t tc_frame_encode_sized(const topos_frame_config* cfg, const topos_frame_input* input,
                              uint32_t target_bytes, uint8_t qp_min, uint8_t qp_max,
                              uint8_t* qp_used, uint8_t* out, size_t out_cap,
                              topos_frame_stats* stats);
/* reference topos_codec.h:336-339: planes_out == NULL is query mode */
int32_t tc_frame_decode(const uint8_t* data, size_t size,
                        uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                        const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info);
/* reference topos_codec.h:398-400: flat count x 4 plane pointers */
int32_t tc_frame_decode_batch(const topos_batch_packet* packets, uint32_t count,
                              uint16_t* const* planes_out, const size_t* strides,
                              topos_frame_output* infos);
/* reference topos_codec.h:410-411 */
int32_t tc_frame_plane_geometry(const topos_frame_output* info, uint32_t plane,
                                uint32_t* width, uint32_t* height);

/* decoder contexts (reference topos_codec.h:515-552) */
int32_t tc_decoder_create(const topos_decoder_config* cfg, tc_decoder** out);
void tc_decoder_destroy(tc_decoder* dec);
int32_t tc_decoder_prepare(tc_decoder* dec, const uint8_t* packet, size_t size,
                           topos_frame_output* info);
int32_t tc_decoder_decode(tc_decoder* dec, const uint8_t* packet, size_t size,
                          const topos_plane_view out[TC_FRAME_MAX_PLANES], topos_frame_output* info);
int32_t tc_decoder_decode_batch(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                                const topos_plane_view* out, topos_frame_output* infos);

/* ---- reduced / scaled preview decode (reference topos_codec.h:341-380,
 * 403-407, 441-497, 529-559) ----
 * A target-size decode reads and checks every slice, reconstructs the frame
 * keeping the coefficients a fixed ratio needs (reduced: scan positions up to
 * 24 / 16 / 10 / 4 for 1/2, 1/3, 1/4, 1/8; scaled: all) and samples each
 * target position from the nearest source sample (floor(t * source /
 * target)); a deblocking stream is filtered on a 4-sample grid when the
 * target is exactly half the source. Output planes: target size, 4:2:2
 * chroma (w+1)/2 wide, RAW phase planes half size. */
typedef enum tc_decode_scale {
    TC_DECODE_SCALE_FULL = 0,
    TC_DECODE_SCALE_HALF = 1,
    TC_DECODE_SCALE_THIRD = 2,
    TC_DECODE_SCALE_QUARTER = 3,
    TC_DECODE_SCALE_EIGHTH = 4
} tc_decode_scale;

/* reference topos_codec.h:362-366: ceil(source / divisor) */
int32_t tc_decode_scale_dimensions(uint32_t source_width, uint32_t source_height, tc_decode_scale scale,
                                   uint32_t* target_width, uint32_t* target_height);
/* reference topos_codec.h:345-349 */
int32_t tc_frame_decode_scaled(const uint8_t* data, size_t size, uint32_t target_width, uint32_t target_height,
                               uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                               const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info);
/* reference topos_codec.h:368-372 */
int32_t tc_frame_decode_reduced(const uint8_t* data, size_t size, tc_decode_scale scale,
                                uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                                const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info);
/* reference topos_codec.h:403-407 */
int32_t tc_frame_decode_batch_reduced(const topos_batch_packet* packets, uint32_t count, tc_decode_scale scale,
                                      uint16_t* const* planes_out, const size_t* strides,
                                      topos_frame_output* infos);

enum { TC_DECODE_MEMORY_CPU = 0u, TC_DECODE_MEMORY_HOST_VISIBLE_GPU = 1u };

typedef enum topos_decode_mode {
    TC_DECODE_MODE_FULL = 0u,
    TC_DECODE_MODE_SCALED = 1u,
    TC_DECODE_MODE_REDUCED = 2u,
    TC_DECODE_MODE_AUTO_2K = 3u    /* sources above 2048 on a side: the first fixed ratio that fits */
} topos_decode_mode;

typedef enum tc_decode_quality {
    TC_DECODE_QUALITY_DEFAULT = 0u,
    TC_DECODE_QUALITY_FAST = 1u,
    TC_DECODE_QUALITY_BALANCED = 2u,
    TC_DECODE_QUALITY_HIGH = 3u
} tc_decode_quality;

/* reference topos_codec.h:465-476 */
typedef struct topos_decode_request {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t mode;          /* topos_decode_mode */
    uint32_t scale;         /* tc_decode_scale; REDUCED only */
    uint32_t target_width;  /* SCALED / AUTO_2K; 0 together with height */
    uint32_t target_height;
    uint32_t quality;       /* tc_decode_quality; checked and kept */
    uint32_t memory_type;   /* TC_DECODE_MEMORY_* */
    uint32_t flags;         /* TC_DECODE_FLAG_* */
    uint32_t reserved[4];   /* must be 0 */
} topos_decode_request;

#define TC_DECODE_FLAG_DROP_ALPHA (1u << 0)

/* reference topos_codec.h:488-493: planes_out == NULL is query mode */
int32_t tc_frame_decode_request(const uint8_t* data, size_t size, const topos_decode_request* request,
                                uint16_t* const planes_out[TC_FRAME_MAX_PLANES],
                                const size_t strides[TC_FRAME_MAX_PLANES], topos_frame_output* out_info);
/* reference topos_codec.h:529-532, 541-545, 555-559 */
int32_t tc_decoder_prepare_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                   const topos_decode_request* request, topos_frame_output* info);
int32_t tc_decoder_decode_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                  const topos_decode_request* request,
                                  const topos_plane_view out[TC_FRAME_MAX_PLANES], topos_frame_output* info);
int32_t tc_decoder_decode_batch_request(tc_decoder* dec, const topos_batch_packet* packets, uint32_t count,
                                        const topos_decode_request* request, const topos_plane_view* out,
                                        topos_frame_output* infos);

/* caller-owned output surface (reference topos_codec.h:495-500); the planes
 * are written during the call only and no pointer is kept */
typedef struct topos_decode_surface {
    uint32_t struct_size;
    uint32_t abi_version;
    uint32_t memory_type;   /* TC_DECODE_MEMORY_* */
    uint32_t reserved;
    topos_plane_view planes[TC_FRAME_MAX_PLANES];
} topos_decode_surface;

/* reference topos_codec.h:521: the context's worker budget (0 = default).
 * Slices of a frame run on the GPU in one launch, so the value is kept and
 * reported but does not change the result or the launch. */
int32_t tc_decoder_set_max_slice_workers(tc_decoder* dec, uint32_t workers);
/* reference topos_codec.h:563-565: tc_decoder_decode into surface->planes
 * after the surface's memory type and struct_size / abi_version are checked */
int32_t tc_decoder_decode_surface(tc_decoder* dec, const uint8_t* packet, size_t size,
                                  const topos_decode_surface* surface, topos_frame_output* info);
/* reference topos_codec.h:568-571: request->memory_type must equal the
 * surface's */
int32_t tc_decoder_decode_surface_request(tc_decoder* dec, const uint8_t* packet, size_t size,
                                          const topos_decode_request* request,
                                          const topos_decode_surface* surface, topos_frame_output* info);

/* ---- packed RGB(A) -> planar conversion (reference topos_codec.h:1016-1102) ----
 * The step before encode in every export: packed 8-bit or 16-bit RGB(A)
 * pixels to YUV 4:2:2 / 4:4:4 or GBR planes of 10 or 12 bits, float32
 * arithmetic in the reference's operation order (no contraction, IEEE
 * division, rint half to even), 4:2:2 chroma a box of the two quantised
 * samples. Runs on the GPU; host buffers in and out. */
typedef enum topos_cvt_format {
    TOPOS_CVT_BGR24 = 0,
    TOPOS_CVT_RGB24 = 1,
    TOPOS_CVT_BGRA32 = 2,
    TOPOS_CVT_RGBA32 = 3,
    TOPOS_CVT_BGR48 = 4,
    TOPOS_CVT_RGB48 = 5,
    TOPOS_CVT_BGRA64 = 6,
    TOPOS_CVT_RGBA64 = 7
} topos_cvt_format;

#define TOPOS_CVT_OUT_YUV422 0u
#define TOPOS_CVT_OUT_YUV444 1u
#define TOPOS_CVT_OUT_GBR 2u

/* reference topos_codec.h:1043-1073 */
typedef struct topos_cvt_params {
    uint32_t struct_size;
    uint32_t format;        /* topos_cvt_format */
    uint32_t width;
    uint32_t height;
    uint32_t out_mode;      /* TOPOS_CVT_OUT_* */
    uint32_t bit_depth;     /* 10 or 12 */
    double kr;              /* forward matrix coefficients (ignored for GBR) */
    double kb;
    uint32_t full_range;    /* 1 full / 0 limited */
    uint32_t alpha_shift;   /* alpha left shift = 16 - declared alpha depth */
    uint32_t alpha_opaque;  /* 1: ignore input alpha, write the opaque value */
    const void* src;        /* packed pixels, height x width x channels */
    uint32_t src_stride;    /* bytes per row; 0 = tight */
    uint16_t* y_out;
    uint32_t y_stride;      /* elements per row; 0 = tight */
    uint16_t* u_out;        /* YUV: the U plane (ceil(w/2) wide for 4:2:2); GBR: the B plane */
    uint32_t u_stride;
    uint16_t* v_out;        /* YUV: the V plane; GBR: the R plane (y_out holds G) */
    uint32_t v_stride;
    uint16_t* a_out;        /* alpha plane or NULL */
    uint32_t a_stride;
    uint32_t max_workers;   /* accepted for compatibility; the GPU splits the work itself */
} topos_cvt_params;

/* reference topos_codec.h:1075-1078 */
int32_t tc_convert_packed_rgb(const topos_cvt_params* params);

/* reference topos_codec.h:1080-1097 */
typedef struct topos_planar_444_to_422_params {
    uint32_t struct_size;
    uint32_t width;         /* 4:4:4 input width */
    uint32_t height;
    uint32_t bit_depth;     /* 10, 12 or 16 (validation only) */
    const uint16_t* u_in;
    size_t u_in_stride;     /* elements; 0 = width */
    const uint16_t* v_in;
    size_t v_in_stride;
    uint16_t* u_out;        /* height x ceil(width / 2) */
    size_t u_out_stride;
    uint16_t* v_out;
    size_t v_out_stride;
    uint32_t max_workers;   /* accepted for compatibility */
} topos_planar_444_to_422_params;

/* reference topos_codec.h:1100-1101: (a + b + 1) >> 1 of horizontal pairs,
 * the edge sample repeated for an odd width */
int32_t tc_convert_yuv444_to_422(const topos_planar_444_to_422_params* params);

/* ---- MOV container, video track (reference topos_codec.h:670-1014) -------
 * Mux: ftyp + mdat + moov, the mdat size patched at finish (the sink needs
 * write and seek_write). Same packets and config give the same bytes as the
 * reference writer. Demux: open parses the moov only and builds the sample
 * index; packets are read and their headers checked against tpcC on demand.
 * Sound and timecode tracks are not written (TC_ERR_NOT_IMPLEMENTED) and are
 * skipped when read. */
#define TC_MOVIE_MAX_SAMPLES 67108864u

/* reference topos_codec.h:674-702; reserved[0..4] declare a sound track
 * (all 0 = none), the only form accepted here */
typedef struct topos_movie_config {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t visible_width;
    uint16_t visible_height;
    uint8_t profile;
    uint8_t pixel_format;
    uint8_t bit_depth;
    uint8_t qmatrix_id;
    uint8_t qp_base;
    int8_t qp_delta_luma;
    int8_t qp_delta_chroma;
    uint8_t alpha_mode;
    uint8_t alpha_bit_depth;
    uint8_t alpha_premultiplied;
    uint8_t color_range;
    uint8_t color_primaries;
    uint8_t color_transfer;
    uint8_t color_matrix;
    uint8_t chroma_siting;
    uint16_t sar_num;
    uint16_t sar_den;
    uint32_t timescale;     /* 0 = 24000 */
    uint32_t reserved[8];
} topos_movie_config;

/* reference topos_codec.h:730-760; reserved[0] = the track's bitstream major
 * (tpcC version), reserved[1] = the tpcD tier (0 = none) */
typedef struct topos_movie_info {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t visible_width;
    uint16_t visible_height;
    uint8_t profile;
    uint8_t pixel_format;
    uint8_t bit_depth;
    uint8_t qmatrix_id;
    uint8_t qp_base;
    int8_t qp_delta_luma;
    int8_t qp_delta_chroma;
    uint8_t alpha_mode;
    uint8_t alpha_bit_depth;
    uint8_t alpha_premultiplied;
    uint8_t color_range;
    uint8_t color_primaries;
    uint8_t color_transfer;
    uint8_t color_matrix;
    uint8_t chroma_siting;
    uint16_t sar_num;
    uint16_t sar_den;
    uint32_t timescale;
    uint32_t sample_count;
    uint32_t faststart;     /* 1 = moov before mdat */
    uint32_t index_bytes;   /* memory of the sample index */
    uint32_t reserved[8];
} topos_movie_info;

/* tpcB alpha budget (reference topos_codec.h:762-780) */
#define TC_ALPHA_BUDGET_RATIO_UNSET 0xFFFFu
#define TC_ALPHA_BUDGET_FLAG_OVERRUN 0x0001u
#define TC_ALPHA_BUDGET_FLAG_AUTHORIZED 0x0002u
#define TC_ALPHA_BUDGET_FLAG_ADAPTED 0x0004u
#define TC_ALPHA_BUDGET_FLAG_RATIO_SATURATED 0x0008u
typedef struct topos_alpha_budget_info {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t target_ratio_bp;
    uint16_t actual_ratio_bp;
    uint16_t max_abs_error;
    uint16_t flags;
    uint32_t frame_count;   /* must equal the sample count when read */
    uint32_t reserved[4];
} topos_alpha_budget_info;

typedef struct topos_mux topos_mux;
typedef struct topos_movie topos_movie;

/* tpcD movie metadata (reference topos_codec.h:784-825) */
#define TC_TIER_NONE 0u
#define TC_TIER_PROXY 1u
#define TC_TIER_LT 2u
#define TC_TIER_STANDARD 3u
#define TC_TIER_HQ 4u
#define TC_TIER_4444 5u
#define TC_TIER_4444XQ 6u
#define TC_TIER_LP 7u
#define TC_TIER_RAW 8u
typedef struct topos_movie_meta {
    uint32_t struct_size;
    uint32_t abi_version;
    uint8_t tier_id;
    uint8_t reserved[3];    /* 0 */
    char vendor[16];        /* NUL-terminated, at most 15 bytes */
    char label[64];         /* NUL-terminated, at most 63 bytes */
} topos_movie_meta;

/* trwm payload, carried opaque (reference topos_codec.h:817-829) */
#define TC_RAW_META_MAX 256u
typedef struct topos_raw_meta {
    uint32_t struct_size;
    uint32_t abi_version;
    uint16_t payload_size;  /* 1..TC_RAW_META_MAX */
    uint16_t reserved;      /* 0 */
    uint8_t payload[TC_RAW_META_MAX];
} topos_raw_meta;

/* reference topos_codec.h:833-836 */
int32_t tc_mux_create(const topos_movie_config* cfg, const topos_io* sink, topos_mux** out);
int32_t tc_mux_add_packet(topos_mux* mux, const uint8_t* packet, size_t size, uint64_t pts_tick, uint32_t dur_tick);
/* reference topos_codec.h:841-848 */
int32_t tc_mux_set_alpha_budget(topos_mux* mux, const topos_alpha_budget_info* info);
int32_t tc_mux_set_movie_meta(topos_mux* mux, const topos_movie_meta* meta);
int32_t tc_mux_set_raw_meta(topos_mux* mux, const topos_raw_meta* meta);
/* reference topos_codec.h:856-874: TC_ERR_NOT_IMPLEMENTED here */
int32_t tc_mux_add_audio(topos_mux* mux, const uint8_t* data, size_t size, uint32_t num_samples);
int32_t tc_mux_set_audio_asc(topos_mux* mux, const uint8_t* asc, size_t size);
int32_t tc_mux_set_timecode(topos_mux* mux, uint32_t hh, uint32_t mm, uint32_t ss, uint32_t ff, uint32_t fps,
                            uint32_t drop_frame);
int32_t tc_mux_set_audio_priming(topos_mux* mux, uint32_t samples);
/* reference topos_codec.h:875-876 */
int32_t tc_mux_finish(topos_mux* mux);
void tc_mux_free(topos_mux* mux);
/* reference topos_codec.h:880 */
int32_t tc_movie_faststart(const topos_io* src, const topos_io* dst);
/* reference topos_codec.h:883-892 */
int32_t tc_movie_open(const topos_io* io, topos_movie** out);
int32_t tc_movie_open_fd(int fd, uint64_t length, topos_movie** out);
int32_t tc_movie_info(const topos_movie* movie, topos_movie_info* info);
int32_t tc_movie_alpha_budget(const topos_movie* movie, topos_alpha_budget_info* info);
/* reference topos_codec.h:968 */
int32_t tc_movie_read_raw_meta(const topos_movie* movie, topos_raw_meta* out);
/* reference topos_codec.h:978-979: buf NULL or cap short -> BUFFER_TOO_SMALL
 * with *need_size */
int32_t tc_movie_packet(const topos_movie* movie, uint32_t sample_index, uint8_t* buf, size_t cap,
                        size_t* need_size);
/* reference topos_codec.h:1000-1003: samples [start, start + count) packed
 * into arena, one io read when they are contiguous in the file */
int32_t tc_movie_packet_batch(const topos_movie* movie, uint32_t start, uint32_t count, uint8_t* arena, size_t cap,
                              size_t* offsets_out, size_t* sizes_out, size_t* need_total);
/* reference topos_codec.h:1004-1014 */
int32_t tc_movie_packet_pts(const topos_movie* movie, uint32_t sample_index, uint64_t* pts_tick, uint32_t* dur_tick);
int32_t tc_movie_packet_sync(const topos_movie* movie, uint32_t sample_index, uint8_t* is_sync);
int32_t tc_movie_prev_sync(const topos_movie* movie, uint32_t sample_index, uint32_t* sync_index);
void tc_movie_close(topos_movie* movie);

#ifdef __cplusplus
}
#endif
#endif
