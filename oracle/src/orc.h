/* ORACLE — test infrastructure only (CPU restatement of the Topos intra
 * frame-coding path used to check the GPU library). Only tests, the smoke
 * check and the benchmark's CPU-baseline leg load it; the product never does.
 *
 * Exported entry points use the orc_ prefix so the oracle can be loaded next
 * to the product library (tc_ prefix) in one process.
 */
#ifndef ORC_H
#define ORC_H

#include <stddef.h>
#include <stdint.h>

/* Same byte layout as the public topos_frame_config (ABI version 2). */
typedef struct orc_cfg {
    uint32_t struct_size, abi_version;
    uint16_t visible_width, visible_height;
    uint8_t profile, pixel_format, bit_depth, qmatrix_id, qp_base;
    int8_t qp_delta_luma, qp_delta_chroma;
    uint8_t slice_rows, alpha_mode, alpha_bit_depth, alpha_premultiplied;
    uint8_t color_range, color_primaries, color_transfer, color_matrix, chroma_siting;
    uint16_t sar_num, sar_den;
    uint32_t reserved[8];
} orc_cfg;

typedef struct orc_hdr {
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
    /* derived geometry */
    uint32_t pcw[4], pch[4], pvw[4], pvh[4], bcols[4], brows[4];
} orc_hdr;

typedef struct orc_slice {
    uint32_t size;
    uint8_t plane;
    uint16_t y0, h;
    uint8_t qpd, k1, k2, k3;
    uint32_t crc;
} orc_slice;

enum {
    ORC_OK = 0, ORC_WARN_CONCEALED = 1, ORC_EINVAL = -1, ORC_ENOMEM = -2,
    ORC_EUNSUP_VERSION = -3, ORC_ELIMIT = -8, ORC_EMALFORMED = -9,
    ORC_ETRUNC = -10, ORC_ECRC = -11, ORC_EBUF = -15, ORC_ENOTIMPL = -16
};

extern const int16_t orc_M[64];
extern const uint32_t orc_E[8];
extern const uint32_t orc_W[64];
extern const uint8_t orc_zigzag[64];
extern const uint32_t orc_qp_scale[2][96];
extern const uint16_t orc_qm[5][2][64];

/* transform / quant (orc.c) */
void orc_fdct(const int32_t x[64], int32_t bit_depth, int32_t F[64]);
void orc_quant(const int32_t F[64], const uint16_t qm[64], uint32_t qp, uint32_t tbl,
               uint32_t bit_depth, int32_t q[64]);
void orc_dequant_idct(const int32_t q[64], const uint16_t qm[64], uint32_t qp, uint32_t tbl,
                      uint32_t bit_depth, int32_t xhat[64]);
uint32_t orc_qstep(uint32_t tbl, uint16_t qm, uint32_t qp);
uint32_t orc_crc32(const uint8_t* p, size_t n);
double orc_log2(double x);

int32_t orc_frame_encode(const orc_cfg* cfg, const uint16_t* const planes[4],
                         const size_t strides[4], uint8_t* out, size_t cap, size_t* size_out);
int32_t orc_frame_decode(const uint8_t* pkt, size_t size, uint16_t* const planes[4],
                         const size_t strides[4], uint8_t* slice_status, uint32_t* concealed);
uint64_t orc_color_blocks(const orc_cfg* cfg);

/* target-size encode: qp search over whole-frame encodes (orc_sized.c) */
int32_t orc_frame_encode_sized(const orc_cfg* cfg, const uint16_t* const planes[4], const size_t strides[4],
                               uint32_t target, uint8_t qp_min, uint8_t qp_max, uint8_t* qp_used,
                               uint8_t* out, size_t cap, size_t* size_out, uint32_t* probes);
void orc_sized_reset(void);
void orc_sized_set_mode(int32_t mode);
void orc_sized_set_hint(int32_t disable, int32_t model_disable);
/* the process-wide seed state (qp, valid, bytes, prev, prev_bytes) */
void orc_sized_get_state(uint32_t s[5]);
void orc_sized_put_state(const uint32_t s[5]);

void orc_deblock_plane(uint16_t* pix, size_t stride, uint32_t w, uint32_t h, uint32_t qp,
                       uint32_t tbl, uint32_t bit_depth);
void orc_deblock_plane_unit(uint16_t* pix, size_t stride, uint32_t w, uint32_t h, uint32_t qp, uint32_t tbl,
                            uint32_t bd, uint32_t unit);
int32_t orc_frame_decode_ex(const uint8_t* pkt, size_t size, uint16_t* const planes[4], const size_t strides[4],
                            uint8_t* slice_status, uint32_t* concealed, int32_t skip_alpha);
/* target-size decode: limit = highest kept scan position (63 all) */
int32_t orc_frame_decode_scaled(const uint8_t* pkt, size_t size, uint32_t tw, uint32_t th, uint32_t limit,
                                int32_t skip_alpha, uint16_t* const planes[4], const size_t strides[4],
                                uint8_t* slice_status, uint32_t* concealed);

#endif
