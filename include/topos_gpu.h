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

/* per-kernel times of the last finished encode / decode launch sequence (ms), measured
 * with HIP events on the launch stream when enabled: fdct, slice_enc, scan,
 * emit, slice_dec, idct, deblock */
void tcx_set_timing(int on);
void tcx_get_timing(float out[7]);

#ifdef __cplusplus
}
#endif
#endif
