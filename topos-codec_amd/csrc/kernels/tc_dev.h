// Device-side helpers shared by the gfx950 kernel translation units.
#ifndef TC_DEV_H
#define TC_DEV_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "tcg.h"
#include "tc_math.h"

// Frozen transform matrix M[u][x] and weights W[u][v] = round(2^32/(E_u E_v)),
// and the zigzag scan (native/src/transform/transform_tables.h:10-36,
// native/src/entropy/scan.h:10-32): bitstream constants.
static __device__ const int8_t kTM[64] = {
    13, 13, 13, 13, 13, 13, 13, 13,   13, 11, 7, 1, -1, -7, -11, -13,
    12, 5, -5, -12, -12, -5, 5, 12,   11, -4, -13, -8, 8, 13, 4, -11,
    9, -9, -9, 9, 9, -9, -9, 9,       8, -13, 4, 11, -11, -4, 13, -8,
    5, -12, 12, -5, -5, 12, -12, 5,   1, -7, 11, -13, 13, -11, 7, -1,
};
static __device__ const uint32_t kTE[8] = { 1352, 680, 676, 740, 648, 740, 676, 680 };
static __device__ const uint32_t kTW[64] = TC_TW_INIT;
static __device__ const uint8_t kZigzag[64] = {
    0, 1, 8, 16, 9, 2, 3, 10, 17, 24, 32, 25, 18, 11, 4, 5,
    12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6, 7, 14, 21, 28,
    35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
    58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63,
};
static __device__ const uint8_t kZigzagInv[64] = {
    0, 1, 5, 6, 14, 15, 27, 28, 2, 4, 7, 13, 16, 26, 29, 42,
    3, 8, 12, 17, 25, 30, 41, 43, 9, 11, 18, 24, 31, 40, 44, 53,
    10, 19, 23, 32, 39, 45, 52, 54, 20, 22, 33, 38, 46, 51, 55, 60,
    21, 34, 37, 47, 50, 56, 59, 61, 35, 36, 48, 49, 57, 58, 62, 63,
};

typedef __attribute__((address_space(1))) const uint32_t gcu32;

__device__ static inline uint32_t wuni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ static inline uint32_t bswap32(uint32_t v)
{
    return (v >> 24) | ((v >> 8) & 0xFF00u) | ((v << 8) & 0xFF0000u) | (v << 24);
}

// ---------------------------------------------------------------------------
// wave-wide CRC32 of a byte range: each lane folds a contiguous chunk with the
// bytewise table, the partial registers are shifted over the bytes that follow
// them (multiplication by x^(8k) mod P) and XOR-reduced across the wave.
// ---------------------------------------------------------------------------
__device__ static inline uint32_t wave_crc32(const uint8_t* p, uint32_t n, const uint32_t* tab,
                                      const uint32_t* x2n)
{
    const uint32_t lane = __lane_id();
    const uint32_t per = (n + 63u) >> 6;
    const uint32_t a = min(n, lane * per), b = min(n, a + per);
    uint32_t c = 0u;
    for (uint32_t i = a; i < b; ++i) { c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8); }
    uint32_t part = a < b ? tc_crc_shift(x2n, c, (uint64_t)(n - b)) : 0u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) { part ^= (uint32_t)__shfl_xor((int)part, off); }
    return part ^ tc_crc_shift(x2n, 0xFFFFFFFFu, n) ^ 0xFFFFFFFFu;
}

// Run-time tables uploaded once by tcg_set_tables.
struct TcgConsts {
    double cost[4097];      // 12 - log2(f), the model-choice cost per symbol
    uint32_t crc[256];      // bytewise CRC32 table (reflected 0xEDB88320)
    uint32_t x2n[32];       // x^(2^k) mod P
    uint32_t x16[256];      // x^(128 j) mod P: shift of a raw CRC over 16 j zero bytes
    tcg_vlc_books vlc;      // V2 canonical VLC books (tcg_set_vlc)
};

// rANS2 symbol pairs of the coefficient blocks [blk0, blk0 + nblk) (64 i32
// each, position 0 the DC in scan and natural order alike), added to
// *counter: one pair per block (DC category, first run) plus one per nonzero
// AC coefficient. Every thread of the workgroup takes part.
__device__ static inline void tc_count_pairs(const int32_t* __restrict__ q, uint32_t blk0, uint32_t nblk,
                                             unsigned long long* counter)
{
    const int4* p = (const int4*)(q + (size_t)blk0 * 64u);
    const uint32_t nchunk = nblk * 16u;
    uint32_t n = 0u;
    for (uint32_t c = threadIdx.x; c < nchunk; c += blockDim.x) {
        const int4 v = p[c];
        n += ((c & 15u) != 0u && v.x != 0 ? 1u : 0u) + (v.y != 0 ? 1u : 0u) + (v.z != 0 ? 1u : 0u) +
             (v.w != 0 ? 1u : 0u);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) { n += (uint32_t)__shfl_xor((int)n, off); }
    if ((threadIdx.x & 63u) == 0u) { atomicAdd(counter, (unsigned long long)n); }
    if (threadIdx.x == 0u) { atomicAdd(counter, (unsigned long long)nblk); }
}

// host-side runtime services (tc_runtime.hip)
#define TCG_GRAPH_KEY 20
#define TCG_GRAPHS_PER_CTX 8
struct tcg_graph_ent {
    uint64_t key[TCG_GRAPH_KEY];
    hipGraphExec_t exec;
    uint64_t used;      // last use (LRU eviction)
};
struct tcg_ctx {
    hipStream_t stream;
    hipEvent_t ev[8];
    int nmarks;     // events recorded by the last launch sequence (0 = timing off)
    int kind;       // 1 encode, 2 decode
    tcg_graph_ent graphs[TCG_GRAPHS_PER_CTX];
    int ngraphs;
    uint64_t clock;
    uint64_t graph_hits, graph_builds;
    void* pin;      // pinned host staging for strided host copies (grow-only)
    size_t pin_cap;
    int eager;      // 1: never capture graphs on this context (no graph allocation)
    unsigned long long* pairs_dev;  // symbol-pair counter of timed sequences (allocated on first use)
    int pairs_valid;                // the last timed sequence counted its pairs
};
const TcgConsts* tcg_consts_dev();
int32_t tcg_check(const char* what);
// records event i of the launch sequence when timing is on
void tcg_mark(tcg_ctx* c, int i);
// the context's symbol-pair counter cleared on its stream, for a sequence
// that recorded timing events; null (nothing to count) otherwise
unsigned long long* tcg_pairs_counter(tcg_ctx* c);
// 1 when launch sequences go through captured graphs (graphs on, timing off)
int tcg_use_graphs();
// compute units of the device
uint32_t tcg_cu_count();
// peak shader clock of the device (kHz, 0 if unknown)
uint32_t tcg_clock_khz();
uint32_t tcg_entropy_capacity_enc(uint32_t wpg, uint32_t max_cols);
uint32_t tcg_entropy_capacity_dec(uint32_t wpg, uint32_t max_cols);
// the cached executable graph for key, or null
hipGraphExec_t tcg_graph_find(tcg_ctx* c, const uint64_t* key);
// stores a graph captured for key (evicting the least recently used one)
int32_t tcg_graph_store(tcg_ctx* c, const uint64_t* key, hipGraph_t g, hipGraphExec_t* out);

// Runs a launch sequence on the context's stream. With graphs in use the
// sequence is captured once per distinct key - every pointer, count and
// launch size its kernels receive - into an executable graph and replayed
// with one hipGraphLaunch from then on; a context keeps its
// TCG_GRAPHS_PER_CTX most recently used graphs. Captured sequences contain
// kernel nodes only (no memset or copy nodes).
template <typename F>
int32_t tcg_run(tcg_ctx* c, const uint64_t* key, F&& launch)
{
    if (!tcg_use_graphs() || c->eager) {
        launch();
        return tcg_check("launch");
    }
    hipGraphExec_t exec = tcg_graph_find(c, key);
    if (exec == nullptr) {
        hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) { return tcg_check("hipStreamBeginCapture"); }
        launch();
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || g == nullptr) { return tcg_check("hipStreamEndCapture"); }
        const int32_t rc = tcg_graph_store(c, key, g, &exec);
        if (rc != 0) { return rc; }
        c->graph_builds++;
    } else {
        c->graph_hits++;
    }
    if (hipGraphLaunch(exec, c->stream) != hipSuccess) { return tcg_check("hipGraphLaunch"); }
    return tcg_check("graph launch");
}

// Phase boundary inside one wavefront's slice work: a workgroup-scope fence
// (waits for the wave's outstanding LDS and memory operations and keeps the
// compiler from moving memory accesses across). The waves of a persistent
// workgroup work on different slices, so they never meet at s_barrier.
__device__ static inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// wave-wide sum and maximum (every lane's value, result wave-uniform) and
// inclusive prefix sum over lanes 0..lane
__device__ static inline uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) { v += (uint32_t)__shfl_xor((int)v, off); }
    return wuni(v);
}

__device__ static inline uint32_t wave_max(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t u = (uint32_t)__shfl_xor((int)v, off);
        v = u > v ? u : v;
    }
    return wuni(v);
}

__device__ static inline uint32_t wave_incl_scan(uint32_t v, uint32_t lane)
{
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = (uint32_t)__shfl_up((int)v, (unsigned)off);
        if (lane >= (uint32_t)off) { v += u; }
    }
    return v;
}

// tc_row_bytes with one symbol per lane: lane s holds symbol s's count (0 at
// and beyond the row's symbol count; the counts of a slice sum below 2^32).
// Byte of a used symbol round(count 255 / total) clamped to [1, 255]; an
// all-zero row becomes byte 1 for symbol 0.
__device__ static inline uint32_t wave_row_byte(uint32_t count, uint32_t lane)
{
    const uint32_t total = wave_sum(count);
    if (total == 0u) { return lane == 0u ? 1u : 0u; }
    if (count == 0u) { return 0u; }
    const uint64_t num = (uint64_t)count * 255u + total / 2u;
    const uint64_t v = (num >> 32) == 0u ? (uint64_t)((uint32_t)num / total) : num / total;
    return v == 0u ? 1u : (v > 255u ? 255u : (uint32_t)v);
}

// tc_model_from_bytes with one symbol per lane: lane s holds symbol s's byte
// (0 at and beyond the row's symbol count). Floor share b 4096 / total (at
// least 1 for a used symbol); while the sum is over 4096, one off the largest
// frequency above 1; while under, one onto the largest remainder of a used
// symbol (ties: the lowest symbol; an adjusted symbol's remainder becomes 0).
// Returns 0 with the lane's frequency and cumulative frequency, or -1 (an
// all-zero row, or no fit); the result is wave-uniform.
__device__ static inline int wave_model_from_bytes(uint32_t b, uint32_t lane, uint32_t& fo, uint32_t& co)
{
    const uint32_t total = wave_sum(b);
    if (total == 0u) { return -1; }
    const uint32_t sc = b * 4096u;
    uint32_t f = b == 0u ? 0u : (sc / total > 1u ? sc / total : 1u);
    uint32_t rem = sc % total;
    uint32_t sum = wave_sum(f);
    while (sum > 4096u) {
        const uint32_t key = f > 1u ? f : 0u;
        const uint32_t mx = wave_max(key);
        if (mx == 0u) { break; }
        const uint32_t big = (uint32_t)__builtin_ctzll(__ballot(key == mx));
        if (lane == big) { f -= 1u; rem = 0u; }
        sum -= 1u;
    }
    while (sum < 4096u) {
        const uint32_t key = b != 0u ? rem + 1u : 0u;
        const uint32_t mx = wave_max(key);
        if (mx == 0u) { break; }
        const uint32_t best = (uint32_t)__builtin_ctzll(__ballot(key == mx));
        if (lane == best) { f += 1u; rem = 0u; }
        sum += 1u;
    }
    if (sum != 4096u) { return -1; }
    fo = f;
    co = wave_incl_scan(f, lane) - f;
    return 0;
}

// This wave's first entry and stride in a persistent launch's work list:
// wave k of the grid takes entries k, k + W, k + 2W, ... (W = waves in the
// grid). The list is sorted longest first, so every wave gets a similar share
// of long and short slices. Both values are wave-uniform (SGPRs): the loop
// over entries is a scalar loop with a single exit.
struct WaveSlot { uint32_t first, stride, wave; };
__device__ static inline WaveSlot wave_slot()
{
    WaveSlot w;
    w.wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t wpg = blockDim.x >> 6;
    w.first = blockIdx.x * wpg + w.wave;
    w.stride = gridDim.x * wpg;
    return w;
}

__device__ static inline const uint32_t* tcg_work_order(const tcg_work* w)
{
    return (const uint32_t*)(w + 1);
}

// f64 cost accumulation in the reference's order with no contraction: the
// model choice compares these sums with strict '<', so every rounding must
// match the unfused x86-64 reference build.
__device__ static inline double cost_add(const TcgConsts* K, double bits, uint32_t cnt, uint32_t f)
{
#pragma clang fp contract(off)
    const double term = (double)cnt * K->cost[f];
    return bits + term;
}
__device__ static inline double cost_add_const(double bits, double c)
{
#pragma clang fp contract(off)
    return bits + c;
}

__device__ static inline uint32_t crc32_bytes_tab(const uint32_t* tab, const uint8_t* p, uint32_t n)
{
    uint32_t c = 0xFFFFFFFFu;
    for (uint32_t i = 0; i < n; ++i) { c = tab[(c ^ p[i]) & 0xFFu] ^ (c >> 8); }
    return c ^ 0xFFFFFFFFu;
}

// bit depth > 12 encoder pre-distortion F' = round(F * 2^32 / (W E_u E_v))
// (native/src/transform/transform.c:125-150)
__device__ static inline int32_t predistort(int32_t F, uint32_t nat)
{
    const int64_t w = (int64_t)kTW[nat] * (int64_t)kTE[nat >> 3] * (int64_t)kTE[nat & 7u];
    const int64_t num = (int64_t)F * ((int64_t)1 << 32);
    if (num >= 0) { return (int32_t)((2 * num + w) / (2 * w)); }
    return -(int32_t)((-2 * num + w) / (2 * w));
}

// alpha plane kernels (tc_alpha.hip)
void tcg_alpha_encode_launch(hipStream_t st, const tcg_eplane* planes, const tcg_eslice* slices,
                             uint32_t nslices, const int32_t* unused, uint8_t* scratch,
                             tcg_eresult* results);
void tcg_alpha_decode_launch(hipStream_t st, const tcg_dplane* planes, const tcg_dslice* slices,
                             uint32_t nslices, int32_t* qbuf, int32_t* status);

// V1 Rice / V2 VLC colour slice kernels (tc_bits.hip)
void tcg_bits_encode_launch(hipStream_t st, const TcgConsts* K, const tcg_eplane* planes,
                            const tcg_eslice* slices, uint32_t nslices, const int32_t* qbuf, uint32_t* dcmbuf,
                            uint8_t* scratch, tcg_eresult* results, uint32_t max_cols);
void tcg_bits_decode_launch(hipStream_t st, const TcgConsts* K, const tcg_dplane* planes,
                            const tcg_dslice* slices, uint32_t nslices, int32_t* qbuf, int32_t* status,
                            uint32_t max_cols);

#endif
