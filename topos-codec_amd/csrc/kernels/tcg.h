/* Internal C ABI between the C11 host layer and the gfx950 HIP kernels.
 *
 * Plain structs and pointers only: the host side is compiled as C11 with gcc
 * and never sees HIP types. Every device pointer is carried as uint64_t so the
 * descriptor tables can be uploaded verbatim.
 *
 * Work units
 *  - encode: one descriptor per (frame, plane) [tcg_eplane] and one per band
 *    slice [tcg_eslice]; all frames of a batch share one flat block index
 *    space so a single launch of each kernel covers the whole batch;
 *  - decode: one descriptor per slice [tcg_dslice] and per (frame, plane)
 *    [tcg_dplane].
 */
#ifndef TCG_H
#define TCG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCG_SLICE_BYTES_PER_BLOCK 512u /* entropy scratch bound per 8x8 block */
#define TCG_SLICE_BYTES_EXTRA 1024u    /* prefix tables + state + slack */
#define TCG_PREFIX_MAX 352u            /* flags + 5x29 + 64 + 5x28 = 350 */
/* V1 Rice / V2 VLC slices: worst case per block is 1008 B (Rice: a 63-bit
 * escape for the DC, each of 63 runs and levels, and the end-of-block run)
 * and 529 B (VLC) */
#define TCG_BITS_BYTES_PER_BLOCK 1024u

/* entropy coding of a colour slice */
#define TCG_MODE_RANS2 0u /* rANS2 with context models (major 7, entropy mode 7) */
#define TCG_MODE_RICE 1u  /* V1 bounded Rice (major 1) */
#define TCG_MODE_VLC 2u   /* V2 canonical VLC, codebook version 1 (major 2, entropy mode 1) */

/* the twelve frozen V2 books, index family * 4 + book (family 0 DC
 * category, 1 run, 2 level category), in the lane layouts the kernels use */
typedef struct tcg_vlc_books {
    uint32_t lim[12][64];  /* lane l = 1..20: end of the length-l codes, left-aligned to 20 bits; else 0 */
    uint32_t info[12][64]; /* lane l: first code of length l << 8 | its index in (length, symbol) order */
    uint32_t sym[12][64];  /* (length, symbol) order position -> symbol */
    uint32_t code[12][64]; /* symbol -> code */
    uint32_t len[12][64];  /* symbol -> code length (0: not in the alphabet) */
} tcg_vlc_books;

/* quantiser context of one (plane, qp): natural-order step Q, half step,
 * dead zone and the constants of the exact divide by Q (tc_qdiv32). */
/* encode transform tile: blocks of one block row per workgroup (the engine
 * sizes the grid with it) */
#ifndef TCG_FQ_BLOCKS
#define TCG_FQ_BLOCKS 32u
#endif
/* decode transform tile: blocks of one block row per IDCT tile (8: one wave,
 * a one-wave workgroup per tile) */
#define TCG_IQ_BLOCKS 8u

typedef struct tcg_qctx {
    double qwt[64];      /* the f64 dequantisation weight Q W 2^-32 of natural position 8 u + v,
                            stored at 8 v + u (column-major, as the decoded coefficients);
                            exact: Q W < 2^40 */
    uint32_t qmul[64];   /* the quantiser divide floor(n / Q) = tc_qdiv32(n, qmul, qsh) */
    uint32_t qsh[64];
    uint32_t Q[64];
    uint32_t half[64];
    uint32_t dz[64];
    uint32_t qlim[64];   /* floor(fclamp / Q): larger |q| dequantise to the clamp */
    uint32_t qlim_min;   /* min over the 64 positions of qlim: |q| <= qlim_min never clamps */
    int32_t fclamp;
    uint32_t bit_depth;
    int32_t mid;
    int32_t maxv;
} tcg_qctx;

typedef struct tcg_eplane {
    uint64_t src;        /* device u16 samples, visible area */
    uint32_t stride;     /* elements */
    uint32_t vis_w, vis_h;
    uint32_t bcols, brows;
    uint32_t blk0;       /* first block (flat batch index) of this plane */
    uint32_t qctx;       /* index into the tcg_qctx table */
    uint32_t kind;       /* 0 colour, 1 alpha */
    uint32_t alpha_shift;/* alpha mode 2: pre-quantisation shift (0 = off) */
    uint32_t coded_w;    /* coded width in samples */
    uint32_t tile0;      /* first transform tile of this plane (ascending) */
    uint32_t tile_m, tile_l; /* transform tile t of the plane sits in block row
                                tc_qdiv32(t, tile_m, tile_l) (= t / tiles per row) */
} tcg_eplane;

typedef struct tcg_eslice {
    uint32_t plane;      /* index into the tcg_eplane table */
    uint32_t frame;      /* frame index in the batch */
    uint32_t y0, h;      /* band in block rows */
    uint32_t blk0, nblk; /* flat block range */
    uint64_t scratch;    /* byte offset of this slice's scratch region */
    uint32_t scratch_cap;
    uint32_t qpd;        /* qp_delta_biased written to the slice header */
    uint32_t hdr_plane;  /* plane number written to the slice header */
    uint32_t mode;       /* TCG_MODE_* of a colour slice */
} tcg_eslice;

/* per-slice encode result */
typedef struct tcg_eresult {
    uint32_t start;      /* payload start, byte offset inside the scratch region */
    uint32_t size;       /* payload bytes */
    int32_t status;      /* 0 ok, else a TC_ERR_* code */
    uint32_t k1, k2, k3; /* slice header k fields (alpha Rice parameters) */
    uint32_t crc;        /* payload CRC32 (written by the emit kernel) */
    uint32_t offset;     /* payload offset inside the frame packet */
    uint32_t aux;        /* alpha mode 2: largest |source - coded| in the band */
    uint32_t pad_;
} tcg_eresult;

typedef struct tcg_eframe {
    uint64_t packet;     /* device output buffer */
    uint32_t cap;        /* its capacity */
    uint32_t s0, s1;     /* slice range */
    uint8_t hdr[56];     /* 53-byte header template (size/CRC filled on device) */
} tcg_eframe;

typedef struct tcg_eframe_out {
    uint32_t size;       /* packet bytes */
    int32_t status;
    uint32_t color_payload, alpha_payload;
    uint32_t color_slices, alpha_slices;
} tcg_eframe_out;

/* decode side */
typedef struct tcg_dslice {
    uint64_t payload;    /* device pointer to slice payload */
    uint32_t size;
    uint32_t crc;        /* expected CRC32 from the slice header */
    uint32_t plane;      /* index into the tcg_dplane table */
    uint32_t y0, h;      /* band in block rows */
    uint32_t blk0;       /* first block (flat) of the band */
    uint32_t qctx;       /* index into the tcg_qctx table */
    uint32_t kind;       /* 0 colour, 1 alpha (Rice/MED) */
    uint32_t k1, k2;     /* slice header k fields: alpha / V1 Rice parameters, V2 books */
    uint32_t dc_max;     /* colour: largest legal mapped DC residual */
    uint32_t lv_max;     /* colour: largest legal mapped AC level */
    uint32_t k3;
    uint32_t mode;       /* TCG_MODE_* of a colour slice */
    uint32_t pad_;
} tcg_dslice;

typedef struct tcg_dplane {
    uint64_t dst;        /* device u16 output, visible area (deblocked in place) */
    uint32_t stride;     /* elements */
    uint32_t vis_w, vis_h;
    uint32_t bcols, brows;
    uint32_t blk0;       /* flat block index of block (0,0) */
    uint32_t row_slice;  /* offset into the row->slice table (one u16 per block row) */
    uint32_t kind;       /* 0 colour, 1 alpha */
    uint32_t alpha_mask; /* alpha mode 2 multiple-of check mask */
    uint32_t fill;       /* concealment value */
    uint32_t maxv;       /* largest sample value */
    uint32_t coded_w;
    uint32_t deblock;    /* 1 = run the output deblocking filter on this plane */
    uint32_t db_alpha;   /* deblock strength alpha (beta and clip derive from it) */
    uint32_t tile0;      /* first transform tile of this plane (ascending) */
    uint32_t limit;      /* highest kept scan position; 0 or >= 63 keep all (reduced decode) */
    uint32_t db_unit;    /* deblock grid in samples (4 or 8; 0 = 8) */
    uint32_t sample;     /* target descriptor of a target-size decode: sample this plane */
    uint32_t tile_m, tile_l; /* transform tile t of the plane sits in block row
                                tc_qdiv32(t, tile_m, tile_l) (= t / tiles per row) */
} tcg_dplane;

/* per (plane of the batch, block row): the quantiser context and the slice
 * that codes the row (decode) */
typedef struct tcg_rowrec {
    uint32_t qctx;
    uint32_t slice;
} tcg_rowrec;

/* packed RGB(A) -> planes: everything one launch needs (device addresses) */
typedef struct tcg_cvt {
    uint64_t src, y, u, v, a;   /* a = 0: no alpha plane */
    uint32_t src_stride;        /* bytes per source row */
    uint32_t y_st, u_st, v_st, a_st;  /* elements per output row */
    uint32_t width, height, cw; /* cw = ceil(width / 2) */
    uint32_t mode;              /* TOPOS_CVT_OUT_* (0 4:2:2, 1 4:4:4, 2 GBR) */
    uint32_t nch, smp;          /* channels, bytes per sample */
    uint32_t ro, go, bo, ao;    /* channel offsets in samples */
    uint32_t alpha_shift, opaque;
    uint32_t vec;               /* set by the launcher: aligned 4-channel rows */
    float scale, kr, kb, kf, du, dv, y_mult, y_off, c_mult, c_off, a_nmax, norm_div;
    float norm_rcp;             /* 1 / norm_div, correctly rounded */
    double rdu, rdv;            /* 1 / du and 1 / dv in binary64, correctly rounded */
} tcg_cvt;

/* 4:4:4 -> 4:2:2 chroma (device addresses, strides in elements) */
typedef struct tcg_p422 {
    uint64_t u_in, v_in, u_out, v_out;
    uint64_t uis, vis, uos, vos;
    uint32_t width, height, cw, pad_;
} tcg_p422;

/* ---- runtime ---- */
/* A context is one default-priority stream plus its timing events; work on
 * different contexts runs concurrently on the device. */
typedef struct tcg_ctx tcg_ctx;

int32_t tcg_init(void);
const char* tcg_last_error(void);
int32_t tcg_ctx_create(tcg_ctx** out);
/* 1: this context never captures launch graphs (graph instantiation allocates) */
void tcg_ctx_set_eager(tcg_ctx* c, int eager);

/* The no-allocation decode's device memory: a file-scope __device__ array of
 * TCG_NOALLOC_BYTES (64 MiB of coefficients = 262,144 8x8 blocks, then
 * descriptors and slice status). Its address on the current device, looked up
 * once per device (at tcg_init for the first device); -2 / -8 on failure. */
#define TCG_NOALLOC_Q_BYTES ((size_t)64 << 20)
#define TCG_NOALLOC_DESC_BYTES ((size_t)3 << 19)
#define TCG_NOALLOC_RES_BYTES ((size_t)1 << 18)
#define TCG_NOALLOC_BYTES (TCG_NOALLOC_Q_BYTES + TCG_NOALLOC_DESC_BYTES + TCG_NOALLOC_RES_BYTES)
#define TCG_MAX_DEVICES 16
int32_t tcg_noalloc_scratch(uint64_t* ptr);
void tcg_ctx_destroy(tcg_ctx* c);
int32_t tcg_malloc(uint64_t* dptr, size_t bytes);
int32_t tcg_free(uint64_t dptr);
int32_t tcg_host_alloc(void** hptr, size_t bytes);
int32_t tcg_host_free(void* hptr);
/* copies are queued on the context's stream; the *_sync forms wait for it */
int32_t tcg_h2d(tcg_ctx* c, uint64_t dst, const void* src, size_t bytes);
int32_t tcg_d2h(tcg_ctx* c, void* dst, uint64_t src, size_t bytes);
int32_t tcg_d2h_sync(tcg_ctx* c, void* dst, uint64_t src, size_t bytes);
int32_t tcg_h2d_2d(tcg_ctx* c, uint64_t dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                   size_t rows);
int32_t tcg_d2h_2d(tcg_ctx* c, void* dst, size_t dpitch, uint64_t src, size_t spitch, size_t width, size_t rows);
int32_t tcg_sync(tcg_ctx* c);
/* 0 when everything queued on the context's stream has finished, 1 while
 * work is pending, negative on a device error (never blocks) */
int32_t tcg_query(tcg_ctx* c);
/* device copy of bytes (multiple of 16) with a plain vector-copy kernel */
int32_t tcg_copy(tcg_ctx* c, uint64_t dst, uint64_t src, size_t bytes);
/* byte-exact device-to-device copy, queued on the context's stream */
int32_t tcg_d2d(tcg_ctx* c, uint64_t dst, uint64_t src, size_t bytes);
/* page-locked host memory (hipHostMalloc), NULL on failure */
void* tcg_pinned_alloc(size_t bytes);
void tcg_pinned_free(void* p);
/* upload the (12 - log2 f) cost table, f = 0..4096 (entry 0 unused), and
 * the CRC32 tables (bytewise table, x^(2^k) mod P) */
int32_t tcg_set_tables(const double* cost4097, const uint32_t* crc256, const uint32_t* x2n32);
/* upload the V2 VLC books */
int32_t tcg_set_vlc(const tcg_vlc_books* books);

/* work list of a persistent entropy launch: the entry count, then `count`
 * slice indices, longest slice first (uploaded with the descriptors) */
typedef struct tcg_work {
    uint32_t count;
    uint32_t pad_[3];
} tcg_work;

/* what a batch holds, for launch shapes: the widest plane (blocks) sizes the
 * entropy kernels' DC-row staging in LDS; kernels for slice kinds the batch
 * lacks are not launched */
#define TCG_HAS_RANS2 1u
#define TCG_HAS_BITS 2u
#define TCG_HAS_ALPHA 4u
typedef struct tcg_batch_info {
    uint32_t max_cols;
    uint32_t kinds;      /* TCG_HAS_* */
    /* persistent rANS2 entropy launch: waves per workgroup (1, 2 or 4) and
     * workgroups (0: enough for one slice per wave) */
    uint32_t ent_wpg;
    uint32_t ent_groups;
    uint32_t ent_count;  /* rANS2 colour slices in the work list */
    /* transform kernels (FDCT + quantisation / dequantisation + IDCT): tiles
     * of 32 blocks per workgroup, 1, 2 or 4 */
    uint32_t tile_tpw;
    tcg_work* work;      /* device work list */
} tcg_batch_info;

/* workgroups of a persistent rANS2 launch that fit on the device at once
 * (all compute units), for a batch whose widest plane is max_cols blocks */
uint32_t tcg_entropy_capacity(int encode, uint32_t wpg, uint32_t max_cols);


/* Deblocking runs one thread per g x g tile centred on a block corner, one
 * workgroup of TCG_DB_WAVES waves per tile row: the waves take the row in runs
 * of TCG_DB_TX - 1 consecutive tiles (a run's first lane only supplies half of
 * its first tile) or, where the plane's alignment rules out 16-byte row
 * chunks, runs of TCG_DB_TX tiles. The grid extent for a w x h plane, packed
 * as (1 << 16 | tile rows); a batch launches the component-wise maximum over
 * its planes. */
#define TCG_DB_TX 64u
#ifndef TCG_DB_WAVES
#define TCG_DB_WAVES 4u
#endif
static inline uint32_t tcg_db_grid(uint32_t w, uint32_t h, uint32_t g)
{
    (void)w;
    return 1u << 16 | (h / g + 1u);
}
static inline uint32_t tcg_db_grid_max(uint32_t a, uint32_t b)
{
    const uint32_t x = (a >> 16) > (b >> 16) ? (a >> 16) : (b >> 16);
    const uint32_t y = (a & 0xFFFFu) > (b & 0xFFFFu) ? (a & 0xFFFFu) : (b & 0xFFFFu);
    return x << 16 | y;
}

/* ---- conversions (queued on the context's stream) ---- */
int32_t tcg_convert_rgb(tcg_ctx* c, const tcg_cvt* k);
/* target-size decode, after tcg_decode on the same stream: every target
 * descriptor with sample != 0 takes, per target sample (tx, ty), the source
 * plane's sample (floor(tx * vw / dw), floor(ty * vh / dh)), then the target
 * deblock pass in place; max_px = largest target plane in samples,
 * max_db_tiles = deblock grid extent of the targets (tcg_db_grid, g = 4) */
int32_t tcg_decode_sample(tcg_ctx* c, const tcg_dplane* src_dev, const tcg_dplane* dst_dev, uint32_t nplanes,
                          uint32_t max_px, uint32_t max_db_tiles);
int32_t tcg_convert_444_422(tcg_ctx* c, const tcg_p422* k);

/* ---- encode ---- */
int32_t tcg_encode(tcg_ctx* c, const tcg_qctx* qctx_dev, uint32_t nplanes, const tcg_eplane* planes_dev,
                   uint32_t nslices, const tcg_eslice* slices_dev, uint32_t nframes,
                   const tcg_eframe* frames_dev, uint64_t qbuf, uint64_t dcmbuf,
                   uint32_t max_plane_tiles, uint64_t scratch, tcg_eresult* results_dev,
                   tcg_eframe_out* frame_out_dev, const tcg_batch_info* info);

/* ---- decode ---- */
/* rowrec_dev: nplanes x max_rows records; max_plane_tiles: largest transform
 * tile count of one plane; max_db_tiles: deblock grid extent over the
 * filtered planes (tcg_db_grid, g = 8) */
int32_t tcg_decode(tcg_ctx* c, const tcg_qctx* qctx_dev, uint32_t nplanes, const tcg_dplane* planes_dev,
                   uint32_t nslices, const tcg_dslice* slices_dev, const tcg_rowrec* rowrec_dev,
                   uint32_t max_rows, uint64_t qbuf, uint32_t max_plane_tiles, int32_t* slice_status_dev,
                   uint32_t max_db_tiles, const tcg_batch_info* info);

/* per-kernel times in milliseconds (HIP events on the launch stream) of the
 * context's last launch sequence; valid after tcg_sync when timing is on */
typedef struct tcg_timing {
    float fdct_ms, slice_enc_ms, scan_ms, emit_ms;
    float slice_dec_ms, idct_ms, deblock_ms;
    float convert_ms;   /* conversion kernel */
    /* rANS2 symbol pairs the sequence coded: one per block (DC category and
     * first run) plus one (level, run) pair per nonzero AC coefficient,
     * counted from the coefficients after the sequence (0 = not counted) */
    uint64_t pairs;
} tcg_timing;
void tcg_set_timing(int on);
/* compute units and peak shader clock (kHz) of the device, 0 before init */
void tcg_device_info(uint32_t* cus, uint32_t* clock_khz);
/* launch sequences through captured HIP graphs (default on; the environment
 * variable TOPOS_HIP_GRAPHS=0 turns them off at start-up) */
void tcg_set_graphs(int on);
void tcg_graph_stats(tcg_ctx* c, uint64_t* hits, uint64_t* builds);
int tcg_timing_on(void);
void tcg_ctx_timing(tcg_ctx* c, tcg_timing* t);

#ifdef __cplusplus
}
#endif
#endif
