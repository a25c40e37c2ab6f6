/* topos_gpu.h — device-resident batch entry points of the MI355X backend.
 *
 * These have no counterpart in the reference C ABI. They run the same frame
 * encode / decode as include/topos_codec.h but leave inputs and outputs in
 * HBM and take many frames (mixed sizes and profiles) per call, so one launch
 * sequence covers a whole batch. Callers: batching front ends and the serving
 * loop. Device pointers are plain 64-bit addresses from tcx_dev_alloc.
 */
#ifndef TOPOS_GPU_H
#define TOPOS_GPU_H

#include <stddef.h>
#include <stdint.h>

#include "topos_codec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 1 when alpha planes are encoded/decoded by this build */
int32_t tcx_alpha_supported(void);
/* initialise the GPU backend now (otherwise done by the first call) */
int32_t tcx_gpu_init(void);

/* device memory; every allocation carries 1 KiB of tail slack so packets can
 * be read through aligned windows */
int32_t tcx_dev_alloc(uint64_t* ptr, size_t bytes);
int32_t tcx_dev_free(uint64_t ptr);
int32_t tcx_h2d(uint64_t dst, const void* src, size_t bytes);
int32_t tcx_d2h(void* dst, uint64_t src, size_t bytes);
/* device-to-device copy with a plain vector-copy kernel (bytes multiple of
 * 16); the calibration point for HBM traffic counters */
int32_t tcx_copy(uint64_t dst, uint64_t src, size_t bytes);

/* Encode n frames in one launch sequence. cfgs[i] per frame; dev_planes and
 * dev_strides are n x 4 (strides in elements, 0 = tight); dev_out[i] receives
 * frame i's packet (capacity out_caps[i]). Per-frame status in rcs[i], packet
 * size in out_sizes[i]; returns the first non-OK status. */
int32_t tcx_encode_batch_dev(const topos_frame_config* cfgs, uint32_t n, const uint64_t* dev_planes,
                             const size_t* dev_strides, const uint64_t* dev_out, const size_t* out_caps,
                             uint32_t* out_sizes, int32_t* rcs, topos_frame_stats* stats);

/* Decode n packets resident in HBM (dev_packets[i], from tcx_dev_alloc) into
 * device planes (n x 4). host_packets[i] is a host copy of the same bytes,
 * used only for the structure parse. Per-frame status in rcs[i]; returns the
 * first non-OK status. */
int32_t tcx_decode_batch_dev(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                             const uint64_t* dev_packets, const uint64_t* dev_planes,
                             const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code running under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same decode as tcx_decode_batch_dev,
 * with every intermediate buffer (coefficients, launch descriptors, slice
 * status) inside one fixed 66 MiB device array of the library and all host
 * staging in fixed arrays: the call itself allocates nothing (the first call
 * in the process creates the stream it runs on). Limits per call: 16 frames,
 * 262,144 8x8 blocks over all planes (one 3840x2160 4:2:2 frame), 4096
 * slices; packets and planes resident in HBM. Beyond a limit the call returns
 * TC_ERR_LIMIT_EXCEEDED before any device work. Calls are serialised inside
 * the library, and the fixed array keeps the last call's contents. */
int32_t tcx_decode_batch_dev_noalloc(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                                     const uint64_t* dev_packets, const uint64_t* dev_planes,
                                     const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs);

/* MOV samples [start, start + count) into HBM and decoded as one batch (no
 * reference counterpart; it joins tc_movie_packet_batch, reference
 * topos_codec.h:1000-1003, with tcx_decode_batch_dev): the packets are read
 * with one io read when contiguous, header-checked, uploaded to dev_arena (a
 * tcx_dev_alloc block of arena_cap >= their total bytes) in one copy, and
 * decoded into dev_planes (count x 4). */
int32_t tcx_movie_decode_batch_dev(const topos_movie* mv, uint32_t start, uint32_t count, uint64_t dev_arena,
                                   size_t arena_cap, const uint64_t* dev_planes, const size_t* dev_strides,
                                   topos_frame_output* infos, int32_t* rcs);

/* device-resident batch decode through one decode request (reduced, scaled,
 * AUTO_2K, DROP_ALPHA): dev_planes receive the request's target geometry */
int32_t tcx_decode_batch_dev_request(uint32_t n, const uint8_t* const* host_packets, const size_t* sizes,
                                     const uint64_t* dev_packets, const uint64_t* dev_planes,
                                     const size_t* dev_strides, topos_frame_output* infos, int32_t* rcs,
                                     const topos_decode_request* request);

/* ---- target-size encode of many frames -----------------------------------
 * Each request runs the search of tc_frame_encode_sized on its own frame
 * (device planes in, device packet out) with its own seed state; the
 * searches of all requests share every launch sequence: each round encodes
 * every unfinished frame at the qp its search needs next and at nearby qps.
 * stream: the request's seed state, updated on success (NULL: no seed). */
typedef struct tcx_sized_stream {
    uint32_t qp, valid, bytes, prev, prev_bytes;
} tcx_sized_stream;

typedef struct tcx_sized_req {
    topos_frame_config cfg;       /* qp_base: where an unseeded search starts */
    uint64_t dev_planes[4];
    size_t dev_strides[4];        /* elements; 0 = visible width */
    uint32_t target_bytes;
    uint8_t qp_min, qp_max;
    tcx_sized_stream* stream;
    uint64_t dev_out;
    size_t dev_out_cap;
    /* results */
    int32_t rc;
    uint8_t qp_used;
    uint32_t packet_size;
    uint32_t encodes;             /* probes + final, counted as the reference counts them */
    topos_frame_stats stats;
} tcx_sized_req;

/* returns the first non-OK request status */
int32_t tcx_encode_sized_batch_dev(tcx_sized_req* reqs, uint32_t n);
/* process-wide seed state of tc_frame_encode_sized: forget it; search mode
 * (0 strict, 1 legacy linear, 2 r6); seed / seed-model switches */
void tcx_sized_reset(void);
void tcx_sized_set_mode(int32_t mode);
void tcx_sized_set_seed(int32_t disable, int32_t model_disable);

/* One serving step: an encode batch and a decode batch run concurrently, each
 * on its own stream (either may be NULL or empty); returns when both are done.
 * Fields mean what the arguments of tcx_encode_batch_dev / tcx_decode_batch_dev
 * mean. Returns the first non-OK status of the encode batch, else of the
 * decode batch. */
typedef struct tcx_enc_batch {
    const topos_frame_config* cfgs;
    uint32_t n;
    const uint64_t* dev_planes;
    const size_t* dev_strides;
    const uint64_t* dev_out;
    const size_t* out_caps;
    uint32_t* out_sizes;
    int32_t* rcs;
    topos_frame_stats* stats;
} tcx_enc_batch;

typedef struct tcx_dec_batch {
    uint32_t n;
    const uint8_t* const* host_packets;
    const size_t* sizes;
    const uint64_t* dev_packets;
    const uint64_t* dev_planes;
    const size_t* dev_strides;
    topos_frame_output* infos;
    int32_t* rcs;
} tcx_dec_batch;

int32_t tcx_serve_step(const tcx_enc_batch* enc, const tcx_dec_batch* dec);

/* the conversions of include/topos_codec.h with every buffer in HBM: the
 * pointers of the parameter struct are device addresses (tcx_dev_alloc);
 * same rules, constants and results, no staging copies */
int32_t tcx_convert_packed_rgb_dev(const topos_cvt_params* params);
int32_t tcx_convert_yuv444_to_422_dev(const topos_planar_444_to_422_params* params);
/* kernel time (ms, HIP events) of the last device-resident conversion while
 * tcx_set_timing is on */
float tcx_last_convert_ms(void);

/* ---- serving loop ------------------------------------------------------
 * A request loop over the batch engine. Encode and decode requests of any
 * size, profile and format are accepted as they arrive (tcx_server_submit
 * never blocks), admitted in arrival order against a budget of HBM, and
 * grouped into launch sequences whenever an engine lane is free (continuous
 * batching): a request that arrives while others run joins the next batch
 * and its work overlaps theirs on the device. Inputs may be resident in HBM
 * or host buffers staged into the budget; results stay in HBM until the
 * caller releases them. All times are tcx_now_ns() values. */
typedef struct tcx_server tcx_server;

typedef struct tcx_server_config {
    size_t hbm_budget;      /* bytes of HBM for admitted requests (staged inputs, packets, planes) */
    uint32_t max_batch;     /* most frames per launch sequence, 1..256 (0: 8) */
    uint32_t max_inflight;  /* launch sequences in flight at once, 1..4 (0: 4) */
    uint32_t gather_us;     /* arrival coalescing: while requests keep arriving less than
                               gather_us apart, a launch waits for them, at most 20 x gather_us
                               after its oldest request arrived, then splits the queue into
                               equal batches (0: launch at once) */
} tcx_server_config;

#define TCX_REQ_ENCODE 1u
#define TCX_REQ_DECODE 2u

typedef struct tcx_request {
    uint32_t kind;                    /* TCX_REQ_ENCODE or TCX_REQ_DECODE */
    uint64_t user;                    /* caller's tag, returned in the completion */
    /* encode: the frame config and its input planes, resident (dev_planes) or
     * host (host_planes, staged into the budget at launch); strides in
     * elements, 0 = tight */
    topos_frame_config cfg;
    uint64_t dev_planes[4];
    const uint16_t* host_planes[4];
    size_t strides[4];
    /* decode: host bytes of the packet (structure parse, and the staged copy
     * unless dev_packet holds a resident one); must stay valid until the
     * request completes */
    const uint8_t* packet;
    size_t packet_size;
    uint64_t dev_packet;
    /* outputs; 0 = allocated by the server from its budget */
    uint64_t dev_out;                 /* encode: packet buffer */
    size_t dev_out_cap;
    uint64_t dev_out_planes[4];       /* decode: output planes */
    size_t out_strides[4];
    topos_frame_output* info;         /* decode: slice status etc., filled by completion (optional) */
    topos_frame_stats* stats;         /* encode: filled by completion (optional) */
} tcx_request;

typedef struct tcx_completion {
    uint64_t id, user;
    uint32_t kind;
    int32_t status;                   /* TC_OK, TC_WARN_CONCEALED or a TC_ERR_* code */
    uint32_t packet_size;             /* encode */
    uint32_t batch_frames;            /* frames in the launch sequence that ran it */
    uint64_t dev_packet;              /* encode: the packet in HBM */
    uint64_t dev_planes[4];           /* decode: the planes in HBM */
    size_t strides[4];                /* decode: their strides (elements) */
    int64_t t_submit_ns, t_admit_ns, t_launch_ns, t_done_ns;
} tcx_completion;

typedef struct tcx_server_stats {
    uint64_t submitted, completed, batches, batch_frames;
    uint64_t budget_bytes, held_bytes, peak_bytes;
    uint32_t queued, inflight;
} tcx_server_stats;

int64_t tcx_now_ns(void);
int32_t tcx_server_create(const tcx_server_config* cfg, tcx_server** out);
/* validates the request (config rules / packet structure) and queues it;
 * a request larger than the whole budget fails with TC_ERR_OUT_OF_MEMORY */
int32_t tcx_server_submit(tcx_server* s, const tcx_request* req, uint64_t* id);
/* waits up to timeout_ns (0: no wait, < 0: forever) for completions and
 * returns up to max of them in completion order (*n may be 0) */
int32_t tcx_server_poll(tcx_server* s, tcx_completion* out, uint32_t max, uint32_t* n, int64_t timeout_ns);
/* every completed request is released once: its server-owned HBM returns
 * to the budget */
int32_t tcx_server_release(tcx_server* s, uint64_t id);
int32_t tcx_server_get_stats(tcx_server* s, tcx_server_stats* st);
/* finishes queued and running work, then frees everything */
void tcx_server_destroy(tcx_server* s);

/* ---- autotuner ----------------------------------------------------------
 * Launch knobs that never change the output: for the persistent rANS2
 * entropy launches, waves (slices) per workgroup, 1, 2 or 4, and the grid,
 * one wave per slice or only the workgroups that fit on the device at once;
 * for the transform kernels, tiles of 32 blocks per workgroup, 1, 2 or 4.
 * Their best values depend on the resolution and profile mix;
 * tcx_autotune_batch times the candidates on a representative batch
 * (arguments as for tcx_encode_batch_dev; the packets it writes are the
 * batch's packets) and keeps the fastest per (kernel, widest plane, slice or
 * tile count class) for the launches that follow. TOPOS_TUNE_FILE names a file the table is loaded
 * from at start-up and saved to after each run. Per-kernel timing is on
 * (graphs off) for the duration of the run. */
typedef struct tcx_tune_report {
    uint32_t candidates;
    uint32_t enc_wpg, enc_persistent, dec_wpg, dec_persistent;  /* chosen shapes */
    float enc_ms_best, enc_ms_default, dec_ms_best, dec_ms_default;  /* entropy kernel times */
    uint32_t fdct_tpw, idct_tpw;  /* chosen transform tiles per workgroup */
    float fdct_ms_best, fdct_ms_default, idct_ms_best, idct_ms_default;
} tcx_tune_report;
int32_t tcx_autotune_batch(const topos_frame_config* cfgs, uint32_t n, const uint64_t* dev_planes,
                           const size_t* dev_strides, const uint64_t* dev_out, const size_t* out_caps, uint32_t reps,
                           tcx_tune_report* report);
/* forgets every tuned shape (launches go back to one wave per slice) */
void tcx_tune_reset(void);
/* fixes the transform tiling of the calling thread's launches: tpw tiles of
 * 32 blocks per workgroup (1, 2 or 4; 0 returns to the tuned tiling) */
void tcx_tune_force_tiles(uint32_t tpw);
/* fixes the rANS2 launch shape of the calling thread's launches: wpg waves
 * per workgroup (1, 2 or 4) and groups workgroups (0: one wave per slice);
 * wpg 0 returns to the tuned shapes */
void tcx_tune_force(uint32_t wpg, uint32_t groups);

/* launch sequences through captured HIP graphs: each distinct batch shape is
 * captured once per engine context and replayed with one graph launch (on by
 * default; TOPOS_HIP_GRAPHS=0 in the environment turns it off at start-up).
 * Per-kernel timing (tcx_set_timing) runs the sequences eagerly. */
void tcx_set_graphs(int on);
/* graph replays and captures so far, summed over the engine contexts */
void tcx_graph_stats(uint64_t* replays, uint64_t* captures);

/* per-kernel times of the last finished encode / decode launch sequence (ms), measured
 * with HIP events on the launch stream when enabled: fdct, slice_enc, scan,
 * emit, slice_dec, idct, deblock */
void tcx_set_timing(int on);
void tcx_get_timing(float out[7]);
/* rANS2 symbol pairs coded by the same two sequences (encode, decode), counted
 * from their coefficients while timing is on: one pair per 8x8 block (DC
 * category, first run) plus one (level, run) pair per nonzero AC coefficient */
void tcx_get_timing_pairs(uint64_t out[2]);
/* compute units and peak shader clock (kHz) of the device the backend runs
 * on (initialises the GPU; both 0 on failure) */
int32_t tcx_device_info(uint32_t* cus, uint32_t* clock_khz);

/* the host-built constant tables the kernels receive, for checking them:
 * tcx_ctx_log2 is the model-choice log2 series (f in 1..4096 gives the
 * per-symbol cost 12 - log2 f); tcx_crc_tables gives the byte table of the
 * reflected CRC32 (0xEDB88320) and x^(2^k) mod P for k = 0..31 */
double tcx_ctx_log2(double x);
void tcx_crc_tables(uint32_t crc256[256], uint32_t x2n32[32]);

#ifdef __cplusplus
}
#endif
#endif
