/* Internal declarations of the C11 host layer (framing rules, error state,
 * device engine). Not installed; the public surface is include/topos_codec.h.
 */
#ifndef TC_HOST_H
#define TC_HOST_H

#include <stddef.h>
#include <stdint.h>

#include "topos_codec.h"
#include "../kernels/tcg.h"

#define TC_HDR_SIZE 53u
#define TC_SLICE_HDR_SIZE 17u
#define TC_MAX_DIM 16384u
#define TC_MAX_SLICE_COUNT 512u
#define TC_MAX_PACKET 268435456u
#define TC_QP_MAX 95u
#define TC_TRANSFER_TRAW_LOG 20u

/* ---- error state (tc_status.c) ---- */
int32_t tc_fail(int32_t code, const char* fmt, ...);
void tc_clear_error(void);

/* ---- frame header model (tc_format.c) ---- */
typedef struct tc_fhdr {
    uint8_t major, minor;
    uint16_t flags;
    uint8_t profile, pixel_format, bit_depth, alpha_mode, alpha_bit_depth;
    uint8_t frame_type;
    uint16_t gop_id;
    uint8_t ref_distance;
    uint16_t coded_w, coded_h, visible_w, visible_h;
    uint8_t plane_count, qmatrix_id, qp_base;
    uint16_t slice_count;
    uint8_t color_range, color_primaries, color_transfer, color_matrix, chroma_siting;
    uint16_t sar_num, sar_den;
    uint32_t packet_size;
    uint8_t entropy_mode, codebook_version, coding_mode;
    /* derived per-plane geometry */
    uint16_t pcw[4], pch[4], pvw[4], pvh[4], bcols[4], brows[4];
} tc_fhdr;

typedef struct tc_shdr {
    uint32_t size;
    uint8_t plane;
    uint16_t y0, h;
    uint8_t qpd, k1, k2, k3;
    uint32_t crc;
} tc_shdr;

typedef struct tc_view {
    tc_fhdr fh;
    uint16_t nslices;
    tc_shdr sl[TC_MAX_SLICE_COUNT];
    uint32_t off[TC_MAX_SLICE_COUNT]; /* payload offset inside the packet */
} tc_view;

extern const uint32_t tc_qp_scale_tab[2][96];
extern const uint16_t tc_qm_tab[5][2][64];
extern const uint8_t tc_zigzag_inv[64];

uint32_t tc_crc32(const uint8_t* p, size_t n);
/* running CRC32: tc_crc32_update(0, ...) over consecutive pieces */
uint32_t tc_crc32_update(uint32_t crc, const uint8_t* p, size_t n);
void tc_crc_tables(uint32_t crc256[256], uint32_t x2n32[32]);
/* V2 VLC books (tc_vlc.c) */
const uint8_t* tc_vlc_lengths(uint32_t family, uint32_t book, uint32_t* nsym);
int32_t tc_vlc_books_build(tcg_vlc_books* b);
double tc_ctx_log2(double x);

int32_t tc_cfg_validate(const topos_frame_config* cfg);
void tc_cfg_to_fhdr(const topos_frame_config* cfg, tc_fhdr* fh);
int32_t tc_fhdr_geometry(tc_fhdr* fh);
int32_t tc_fhdr_validate(const tc_fhdr* fh);
void tc_fhdr_pack(const tc_fhdr* fh, uint8_t out[TC_HDR_SIZE]);
int32_t tc_fhdr_parse(const uint8_t* data, size_t size, tc_fhdr* fh);
int32_t tc_band_plan(const tc_fhdr* fh, uint32_t slice_rows, uint32_t bands[4], uint32_t* total);
int32_t tc_parse_structure(const uint8_t* data, size_t size, tc_view* v);
uint32_t tc_slice_rows_of(const topos_frame_config* cfg);
int32_t tc_qp_eff(const tc_fhdr* fh, const topos_frame_config* cfg, uint32_t plane);
int tc_is_chroma(const tc_fhdr* fh, uint32_t p);
int tc_is_alpha(const tc_fhdr* fh, uint32_t p);
void tc_qctx_build(tcg_qctx* q, const uint16_t* qm, uint32_t qp, uint32_t bit_depth, uint32_t tbl);
uint32_t tc_deblock_alpha(uint32_t tbl, uint32_t qp);
void tc_fill_output_info(const tc_fhdr* fh, topos_frame_output* info);

/* decode-side checks (tc_api.c): tc_dec_prepare parses a packet's structure
 * and stamps info (a failure rejects the packet); tc_dec_supported says
 * whether the frame's bitstream generation decodes here; _check is both */
int32_t tc_dec_prepare(const uint8_t* data, size_t size, topos_frame_output* info, tc_view* v);
int32_t tc_dec_supported(const tc_view* v);
int32_t tc_dec_prepare_check(const uint8_t* data, size_t size, topos_frame_output* info, tc_view* v);

/* request rules of topos_decode_request and the batch decode core
 * (tc_preview.c); scale is a tc_decode_scale used when request == NULL */
int32_t tc_decode_request_validate(const topos_decode_request* q);
int32_t tc_decode_batch_run(const topos_batch_packet* packets, uint32_t count, uint32_t scale,
                            const topos_decode_request* request, uint16_t* const* planes_out,
                            const size_t* strides, topos_frame_output* infos);

/* encode argument checks of tc_frame_encode (tc_api.c): host planes or device planes */
int32_t tc_enc_check_host(const topos_frame_config* cfg, const topos_frame_input* input);
int32_t tc_enc_check_dev(const topos_frame_config* cfg, const uint64_t* dev_planes, const size_t* dev_strides);

/* ---- device engine (tc_engine.c) ---- */
typedef struct tc_ectx tc_ectx;
/* lazily initialises the GPU; returns a free context (locked until release) */
int32_t tc_engine_acquire(tc_ectx** out);
/* never blocks: 1 (and *out = NULL) when every context is in use */
int32_t tc_engine_try_acquire(tc_ectx** out);
void tc_engine_release(tc_ectx* c);
/* the no-allocation context (locked until release): fixed device buffers in
 * the static scratch, fixed host staging, eager launches; plans beyond
 * TC_NA_MAX_FRAMES frames / TC_NA_MAX_SLICES slices or its buffers fail with
 * TC_ERR_LIMIT_EXCEEDED */
#define TC_NA_MAX_FRAMES 16u
#define TC_NA_MAX_SLICES 4096u
int32_t tc_engine_acquire_noalloc(tc_ectx** out);
/* the context's input / output staging buffers, grown to at least the sizes
 * given (conversions of host buffers) */
int32_t tc_engine_staging(tc_ectx* c, size_t in_bytes, size_t out_bytes, uint64_t* in, uint64_t* out);
/* device packet capacity the encoder needs for one frame of cfg (0 on a bad
 * config): frame header plus every slice's header and entropy bound */
size_t tc_engine_encode_cap(const topos_frame_config* cfg);
tcg_ctx* tc_engine_stream(tc_ectx* c);
/* launch shape of a persistent rANS2 launch (tc_tune.c): waves per
 * workgroup and workgroups (0 = one wave per slice) for a batch whose widest
 * colour plane is cols blocks and which holds count rANS2 slices */
void tc_tune_entropy(int encode, uint32_t cols, uint32_t count, uint32_t nframes, uint32_t* wpg, uint32_t* groups);
/* transform tiles per workgroup (1, 2, 4) for a batch of `tiles` 32-block
 * transform tiles */
uint32_t tc_tune_tiles(int encode, uint32_t cols, uint32_t tiles);
/* graph replays / captures summed over the contexts */
void tc_engine_graph_stats(uint64_t* hits, uint64_t* builds);
/* per-kernel times of the last finished encode and decode (timing on) */
void tc_engine_last_timing(tcg_timing* enc, tcg_timing* dec);
/* keeps the conversion kernel time of the context's last launch (timing on) */
void tc_engine_record_convert(tc_ectx* c);
float tc_engine_last_convert_ms(void);

/* encode n frames (all with the same cfg-derived header but separate inputs)
 * into device packets; host_out[i] receives frame i's packet when non-NULL */
typedef struct tc_enc_req {
    const topos_frame_config* cfg;
    const topos_frame_input* input;   /* host planes (API path) */
    const uint64_t* dev_planes;       /* device planes (resident path), or NULL */
    const size_t* dev_strides;        /* elements */
    uint8_t* out;                     /* host output (may be NULL) */
    size_t out_cap;
    uint64_t dev_out;                 /* device output packet (resident path) */
    size_t dev_out_cap;
    topos_frame_stats* stats;
    int32_t rc;
    uint32_t packet_size;
} tc_enc_req;

int32_t tc_engine_encode(tc_ectx* c, tc_enc_req* reqs, uint32_t n);
/* split form: submit queues everything on the context's stream; finish waits
 * and fills the requests (the requests and inputs must stay valid between) */
int32_t tc_engine_encode_submit(tc_ectx* c, tc_enc_req* reqs, uint32_t n);
int32_t tc_engine_encode_finish(tc_ectx* c, tc_enc_req* reqs, uint32_t n);

typedef struct tc_dec_req {
    const uint8_t* data;              /* host packet (API path) */
    size_t size;
    uint64_t dev_data;                /* device packet (resident path), or 0 */
    uint16_t* const* planes;          /* host planes (4), or NULL */
    const size_t* strides;            /* host strides (4), or NULL */
    const uint64_t* dev_planes;       /* device planes (resident path), or NULL */
    const size_t* dev_strides;
    topos_frame_output* info;
    const tc_view* view;              /* parsed structure */
    /* target-size decode (0 = source size): output planes are tw x th (4:2:2
     * chroma and RAW planes as the frame geometry derives them); limit =
     * highest kept scan position (63 all); skip_alpha drops plane 3 */
    uint32_t tw, th, limit, skip_alpha;
    int32_t rc;
} tc_dec_req;

/* output plane size of plane p for a target-size decode */
void tc_target_plane_dims(const tc_fhdr* fh, uint32_t tw, uint32_t th, uint32_t p, uint32_t* w, uint32_t* h);

int32_t tc_engine_decode(tc_ectx* c, tc_dec_req* reqs, uint32_t n);
int32_t tc_engine_decode_submit(tc_ectx* c, tc_dec_req* reqs, uint32_t n);
int32_t tc_engine_decode_finish(tc_ectx* c, tc_dec_req* reqs, uint32_t n);

#endif
