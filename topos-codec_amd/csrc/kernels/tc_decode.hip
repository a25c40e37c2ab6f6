// gfx950 decode kernels: per-slice CRC check + rANS2 decode (one wavefront per
// slice), fused dequantisation + 8x8 inverse transform + clip-store, and the
// output deblocking filter.
//
// Reference semantics (paths relative to the reference tree):
//   slice decode    native/src/bitstream/slice_codec.c:554-736, codec/codec.c:8204-8494
//   reconstruction  native/src/codec/color_store.h:454-718, transform/transform.c:69-117
//   deblock         native/src/codec/deblock.c:71-334
#include <hip/hip_runtime.h>

#include <type_traits>

#include "tcg.h"
#include "tc_math.h"
#include "tc_dev.h"

// ---------------------------------------------------------------------------
// K5: rANS2 colour slice decoder, one wavefront per slice.
// The decode chain is serial (one 32-bit state per slice) and a wavefront
// issues at most one instruction per four cycles, so the loop is written as a
// short, nearly branch-free chain of wave-uniform (SGPR) arithmetic:
//  - symbol search: each context row is held in two vector registers with
//    symbol s in lane 63 - s: the cumulative frequency (unused symbols hold
//    0xFFFF) and the entry f << 19 | s << 12 | cum. The lowest lane whose
//    cumulative frequency is <= slot (s_ff1 of a compare mask) is the symbol;
//  - renormalisation: the next payload bits wait MSB-first in a 64-bit
//    scalar reservoir; taking k = 0..3 bytes is one 64-bit shift of
//    (state : reservoir) by 8k with k from the state's leading zeros;
//  - the reservoir is refilled with whole big-endian dwords from a 256-byte
//    window held one dword per lane (the following window prefetched), at two
//    points per (run, level) pair, which bounds what a pair can consume;
//  - a run symbol of 63 (end of block) and a position overflow leave the pair
//    loop through one compare; range errors are collected per block.
// The payload must be readable 768 bytes past its end (window prefetch).
// ---------------------------------------------------------------------------

struct RansRd {
    uint32_t x;       // rANS state
    uint64_t R;       // reservoir: next payload bits, MSB-first, rc valid
    uint32_t rc;
    uint32_t widx;    // next dword of `cur` to move into the reservoir
    uint32_t nwin;    // windows fetched so far
    uint32_t maxwin;  // windows that may be fetched (payload end + 256)
    uint32_t loaded;  // payload bits moved into the reservoir
    uint32_t cur;     // lane i: dword i of the current window, big-endian
    uint32_t nxt;     // lane i: dword i of the next window, as stored
    gcu32* wp;        // address of the window after `nxt`
};

__device__ static inline void rr_refill(RansRd& D, uint32_t lane)
{
    if (D.widx == 64u) {
        // past the last window that may hold payload bytes the same window is
        // read again: those bits are only consumed by a malformed stream,
        // which the final length check rejects
        D.cur = bswap32(D.nxt);
        D.nxt = D.wp[lane];
        if (D.nwin < D.maxwin) { D.wp += 64; }
        D.nwin += 1u;
        D.widx = 0u;
    }
    const uint32_t w = __builtin_amdgcn_readlane(D.cur, D.widx);
    D.R |= (uint64_t)w << (32u - D.rc);
    D.rc += 32u;
    D.loaded += 32u;
    D.widx += 1u;
}

__device__ static inline void rr_need(RansRd& D, uint32_t bits, uint32_t lane)
{
    if (__builtin_expect(D.rc < bits, 0)) { rr_refill(D, lane); }
}

// payload = first byte after the 4-byte initial state, n bytes
__device__ static inline void rr_init(RansRd& D, uint32_t x0, const uint8_t* payload, uint32_t n, uint32_t lane)
{
    const uintptr_t a = (uintptr_t)payload;
    const uintptr_t w0 = a & ~(uintptr_t)255u;
    gcu32* wp = (gcu32*)w0;
    D.x = x0;
    D.cur = bswap32(wp[lane]);
    D.nxt = wp[64u + lane];
    D.wp = wp + 128;
    D.nwin = 2u;
    D.maxwin = (uint32_t)((a + n + 256u - w0 + 255u) >> 8);
    const uint32_t off = (uint32_t)(a - w0);
    const uint32_t mis = (off & 3u) * 8u;
    D.widx = off >> 2;
    const uint32_t w = __builtin_amdgcn_readlane(D.cur, D.widx);
    D.R = ((uint64_t)w << 32) << mis;
    D.rc = 32u - mis;
    D.loaded = 32u - mis;
    D.widx += 1u;
}

typedef uint32_t tc_u32x2 __attribute__((ext_vector_type(2)));

// renormalisation by n8 = 8k bits: k bytes move from the reservoir into the state
__device__ static inline void rr_shift(RansRd& D, uint32_t n8)
{
    const tc_u32x2 pair = { (uint32_t)(D.R >> 32), D.x };  // (state : next 32 payload bits)
    const uint64_t t = __builtin_bit_cast(uint64_t, pair);
    D.x = (uint32_t)((t << n8) >> 32);
    D.R <<= n8;
    D.rc -= n8;
}

// bytes to take for a state x > 0: k = ceil((23 - floor(log2 x)) / 8) below
// 2^23, as a bit count
__device__ static inline uint32_t norm_bits(uint32_t x) { return ((uint32_t)__builtin_clz(x) - 1u) & 0x18u; }

__device__ static inline void rr_norm(RansRd& D) { rr_shift(D, norm_bits(D.x)); }

// Row vectors of one context model, symbol s in lane 63 - s: the cumulative
// frequency (0xFFFF for an unused symbol: never <= slot) and the frequency.
// Per lane, for the DC and level rows (symbol = magnitude category c = 63 -
// lane): w = max(c - 1, 0) raw suffix bits, the implicit top bit (c != 0) << w
// when w <= 23, and the previous-category bucket pb the symbol selects; for
// the run row the scan-position step run + 1 (64 = end of block).
struct RowV { uint32_t cum, f; };
struct LaneSym { uint32_t w, topb, pb, step; };

// Decodes one run symbol: every lane forms its symbol's successor state
// f (x >> 12) + slot - cum in one 24-bit multiply-add while the wave compares
// the slot against the cumulative frequencies; the lowest matching lane is the
// symbol, readlanes fetch its state and step, and the scalar unit
// renormalises.
__device__ static inline uint32_t drun(RansRd& D, const RowV& r, uint32_t step)
{
    const uint32_t slot = D.x & 4095u;
    const uint32_t cand = __umul24(r.f, D.x >> 12) + (slot - r.cum);
    const uint32_t L = (uint32_t)__builtin_ctzll(__ballot(r.cum <= slot));
    D.x = __builtin_amdgcn_readlane(cand, L);
    const uint32_t e = __builtin_amdgcn_readlane(step, L);
    rr_norm(D);   // every successor is >= 2^11 > 0
    return e;
}

// lane idx of v becomes m (a scalar): one lane write
__device__ static inline uint32_t set_lane(uint32_t v, uint32_t m, uint32_t idx)
{
    // the lane select goes through M0 (one scalar operand per vector instruction)
    asm("s_mov_b32 m0, %2\n\ts_nop 0\n\tv_writelane_b32 %0, %1, m0" : "+v"(v) : "s"(m), "s"(idx) : "m0");
    return v;
}

__device__ static inline uint32_t rr_bits(RansRd& D, uint32_t w)
{
    const uint32_t c = D.x & ((1u << w) - 1u);
    D.x >>= w;
    rr_norm(D);
    return c;
}

// Decodes one DC or level symbol and its raw suffix; returns the mapped
// magnitude (implicit top bit and suffix) and the symbol's lane. Besides the
// successor state every lane renormalises it from the top reservoir word
// (<= 16 bits) and splits off its own suffix bits, so after the search the
// scalar unit only renormalises once more after the suffix. A suffix over 23
// bits (category above 24: lanes below 39) is taken in two steps, the low 16
// bits first. The reservoir holds >= 16 bits on entry.
__device__ static inline uint32_t dlevel(RansRd& D, const RowV& r, const LaneSym& ls, uint32_t lane, uint32_t& L)
{
    const uint32_t slot = D.x & 4095u;
    const uint32_t cand = __umul24(r.f, D.x >> 12) + (slot - r.cum);
    const uint32_t n8 = norm_bits(cand);
    L = (uint32_t)__builtin_ctzll(__ballot(r.cum <= slot));
    // a 32-bit scalar from here on (the compiler would otherwise test the
    // 64-bit bit count with a vector compare)
    asm("" : "+s"(L));
    const tc_u32x2 pr = { (uint32_t)(D.R >> 32), cand };  // (state : next 32 payload bits)
    const uint32_t x1 = (uint32_t)((__builtin_bit_cast(uint64_t, pr) << n8) >> 32);
    const uint32_t m = ls.topb | __builtin_amdgcn_ubfe(x1, 0u, ls.w);
    const uint32_t xs = x1 >> ls.w;
    const uint32_t n8s = __builtin_amdgcn_readlane(n8, L);
    D.R <<= n8s;
    D.rc -= n8s;
    rr_need(D, 24u, lane);
    if (__builtin_expect(L < 39u, 0)) {
        const uint32_t w = 62u - L;
        D.x = __builtin_amdgcn_readlane(x1, L);
        const uint32_t lo = rr_bits(D, 16u);
        rr_need(D, 16u, lane);
        const uint32_t hi = rr_bits(D, w - 16u);
        return (1u << w) | (hi << 16) | lo;
    }
    D.x = __builtin_amdgcn_readlane(xs, L);
    rr_norm(D);
    return __builtin_amdgcn_readlane(m, L);
}

// the AC pairs of one block, level rows fixed for scan positions below LIM;
// val holds mapped magnitudes by scan position (lane = position)
template <uint32_t LIM>
__device__ static inline void ac_phase(RansRd& D, const RowV& lvl, const RowV& run, const LaneSym& ls,
                                       uint32_t& idx, uint32_t& e, uint32_t& val, uint32_t lane)
{
    while (idx < LIM) {
        uint32_t L;
        const uint32_t m = dlevel(D, lvl, ls, lane, L);
        val = set_lane(val, m, idx);
        rr_need(D, 32u, lane);
        e = drun(D, run, ls.step);
        idx += e;
    }
}

// row r of the per-lane row arrays (r is wave-uniform)
#define ROW(r) RowV{ cumv[r], fv[r] }

template <uint32_t LM, uint32_t DM>
__device__ static int32_t slice_decode_body(RansRd& D, const uint32_t (&cumv)[11], const uint32_t (&fv)[11],
                                            const LaneSym& ls, const tcg_dslice& S, uint32_t cols,
                                            int32_t* __restrict__ qbuf, int32_t* sDcRow, uint32_t lane)
{
    constexpr uint32_t DC_N = DM ? 5u : 1u;
    constexpr uint32_t RUN_ROW = DC_N;
    constexpr uint32_t LVL0 = DC_N + 1u;
    const uint32_t dc_max = S.dc_max, lv_max = S.lv_max;
    const int32_t dlim = dc_max > 0x08000000u ? (1 << 27) : (1 << 25);
    const RowV run = ROW(RUN_ROW);
    // lane = scan position; its coefficient's place in the block, stored
    // column-major (natural index 8 r + c at 8 c + r; blocks of the band are
    // consecutive)
    const uint32_t nat = kZigzag[lane];
    int32_t* out = qbuf + (size_t)S.blk0 * 64u + ((nat & 7u) << 3 | nat >> 3);
    uint32_t pdc_b = 0u;
    int32_t left = 0;
    for (uint32_t by = 0; by < S.h; ++by) {
        for (uint32_t bx = 0; bx < cols; ++bx) {
            const int32_t top_v = sDcRow[bx];
            uint32_t drow = 0u;
            if (DM) { drow = (by == 0u && bx == 0u) ? 0u : 1u + pdc_b; }
            rr_need(D, 32u, lane);
            uint32_t Ld;
            const uint32_t dm = dlevel(D, ROW(drow), ls, lane, Ld);
            const int32_t top = by ? (int32_t)wuni((uint32_t)top_v) : 0;
            // tc_dc_predict in 32 bits: accepted DC values lie within +-dlim
            // <= 2^27, so left + top + 1 cannot overflow, and the rounded
            // mean (s + 1) >> 1 (arithmetic) equals its s >= 0 / s < 0 split;
            // with dm <= dc_max < 2^28 the residual is within +-2^27, so dv is
            // exact whenever the block is accepted
            const int32_t pred = bx != 0u ? (by != 0u ? (left + top + 1) >> 1 : left) : (by != 0u ? top : 0);
            const int32_t dv = pred + (int32_t)((dm >> 1) ^ (0u - (dm & 1u)));   // tc_unmap_signed
            uint32_t bad = (dm > dc_max || dv > dlim || dv < -dlim) ? 1u : 0u;
            const int32_t dc = dv;
            if (lane == 0u) { sDcRow[bx] = dc; }
            left = dc;
            if (DM) { pdc_b = __builtin_amdgcn_readlane(ls.pb, Ld); }
            uint32_t val = 0u;
            rr_need(D, 32u, lane);
            uint32_t e = drun(D, run, ls.step);
            uint32_t idx = e;
            if (LM == 1u) {
                // position buckets (scan position < 8, < 16, < 32, else): the
                // position only grows, so each bucket's row is fixed in its loop
                ac_phase<8u>(D, ROW(LVL0 + 0u), run, ls, idx, e, val, lane);
                ac_phase<16u>(D, ROW(LVL0 + 1u), run, ls, idx, e, val, lane);
                ac_phase<32u>(D, ROW(LVL0 + 2u), run, ls, idx, e, val, lane);
                ac_phase<64u>(D, ROW(LVL0 + 3u), run, ls, idx, e, val, lane);
            } else if (LM == 0u) {
                ac_phase<64u>(D, ROW(LVL0), run, ls, idx, e, val, lane);
            } else {
                // previous-category buckets: the row follows the last level
                uint32_t lrow = LVL0 + 4u;
                while (idx < 64u) {
                    uint32_t L;
                    const uint32_t m = dlevel(D, ROW(lrow), ls, lane, L);
                    lrow = LVL0 + __builtin_amdgcn_readlane(ls.pb, L);
                    val = set_lane(val, m, idx);
                    rr_need(D, 32u, lane);
                    e = drun(D, run, ls.step);
                    idx += e;
                }
            }
            // the loop ends on a position past 63: legal only through the
            // end-of-block run; levels above the legal range reject the slice
            bad |= e != 64u ? 1u : 0u;
            bad |= __ballot(lane != 0u && val > lv_max) != 0ull ? 1u : 0u;
            if (bad) { return -9; }
            const int32_t v = lane == 0u ? dc : (int32_t)((val >> 1) ^ (0u - (val & 1u)));  // tc_unmap_signed
            *out = v;
            out += 64;
        }
    }
    return 0;
}

// LDS words of one decoding wave: model frequencies (11 x 64 u16) and
// cumulative frequencies (11 x 65 u16), then the DC values of the block row
// above (one per block column of the batch's widest plane)
#define DEC_FIXED_WORDS 710u

__host__ __device__ static inline uint32_t dec_wave_words(uint32_t cols) { return (DEC_FIXED_WORDS + cols + 1u) & ~1u; }

__device__ static void slice_decode_one(const TcgConsts* __restrict__ K, const tcg_dplane* __restrict__ planes,
                                        const tcg_dslice* __restrict__ slices, int32_t* __restrict__ qbuf,
                                        int32_t* __restrict__ status, uint32_t si, uint32_t lane,
                                        const uint32_t* sTab, const uint32_t* sX2n, uint32_t* sW)
{
    uint16_t (*const sFreq)[64] = (uint16_t (*)[64])sW;
    uint16_t (*const sCum)[65] = (uint16_t (*)[65])(sW + 352);
    int32_t* const sDcRow = (int32_t*)(sW + DEC_FIXED_WORDS);

    const tcg_dslice S = slices[si];
    if (S.kind != 0u || S.mode != TCG_MODE_RANS2) { return; }
    const tcg_dplane P = planes[S.plane];
    wave_sync();

    const uint8_t* pl = (const uint8_t*)S.payload;
    const uint32_t n = S.size;
    int32_t rc = 0;
    if (wuni(wave_crc32(pl, n, sTab, sX2n)) != S.crc) { rc = -11; }
    uint32_t flags = 0u, lvl_model = 0u, dc_model = 0u, lvl_n = 1u, dc_n = 1u, prefix = 0u;
    if (rc == 0 && n < 1u + 121u + 4u) { rc = -10; }
    if (rc == 0) {
        flags = wuni(pl[0]);
        lvl_model = flags & 3u;
        dc_model = (flags >> 2) & 1u;
        if ((flags & 0xF8u) != 0u || lvl_model > 2u) { rc = -9; }
        lvl_n = lvl_model == 1u ? 4u : (lvl_model == 2u ? 5u : 1u);
        dc_n = dc_model ? 5u : 1u;
        prefix = 1u + dc_n * 29u + 64u + lvl_n * 28u;
        if (rc == 0 && n < prefix + 4u) { rc = -10; }
    }
    const uint32_t rows_n = dc_n + 1u + lvl_n;
    if (rc == 0) {
        // rows: [0, dc_n) dc, dc_n run, then lvl rows; one row at a time, one
        // symbol per lane
        for (uint32_t r = 0; r < rows_n; ++r) {
            uint32_t off, nsym;
            if (r < dc_n) { off = 1u + r * 29u; nsym = 29u; }
            else if (r == dc_n) { off = 1u + dc_n * 29u; nsym = 64u; }
            else { off = 1u + dc_n * 29u + 64u + (r - dc_n - 1u) * 28u; nsym = 28u; }
            const uint32_t b = lane < nsym ? (uint32_t)pl[off + lane] : 0u;
            uint32_t f = 0u, c = 0u;
            if (wave_model_from_bytes(b, lane, f, c) != 0) { rc = -9; break; }
            if (lane < nsym) { sFreq[r][lane] = (uint16_t)f; sCum[r][lane] = (uint16_t)c; }
            if (lane == 0u) { sCum[r][nsym] = 4096u; }
        }
        wave_sync();
    }
    if (rc == 0) {
        // symbol s of row r sits in lane 63 - s
        uint32_t cumv[11], fv[11];
        const uint32_t sym = 63u - lane;
        const uint32_t w = sym > 1u ? sym - 1u : 0u, nz = sym != 0u ? 1u : 0u;
        LaneSym ls;
        ls.w = w;
        ls.topb = (nz && w <= 23u) ? (1u << w) : 0u;
        ls.pb = tc_ctx_prev_wnz(w, nz);
        ls.step = sym + 1u;
#pragma unroll
        for (uint32_t r = 0; r < 11u; ++r) {
            const uint32_t nsym = r < dc_n ? 29u : (r == dc_n ? 64u : 28u);
            const bool live = r < rows_n && sym < nsym;
            cumv[r] = live ? (uint32_t)sCum[r][sym] : 0xFFFFu;
            fv[r] = live ? (uint32_t)sFreq[r][sym] : 0u;
        }
        const uint32_t x0 = wuni(((uint32_t)pl[prefix] << 24) | ((uint32_t)pl[prefix + 1u] << 16) |
                                 ((uint32_t)pl[prefix + 2u] << 8) | (uint32_t)pl[prefix + 3u]);
        if (x0 < (1u << 23)) { rc = -9; }
        if (rc == 0) {
            RansRd D;
            rr_init(D, x0, pl + prefix + 4u, n - prefix - 4u, lane);
            const uint32_t cols = P.bcols;
            switch (flags & 7u) {
            case 0u: rc = slice_decode_body<0u, 0u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            case 1u: rc = slice_decode_body<1u, 0u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            case 2u: rc = slice_decode_body<2u, 0u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            case 4u: rc = slice_decode_body<0u, 1u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            case 5u: rc = slice_decode_body<1u, 1u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            default: rc = slice_decode_body<2u, 1u>(D, cumv, fv, ls, S, cols, qbuf, sDcRow, lane); break;
            }
            // the stream must be consumed exactly
            if (rc == 0 && D.loaded - D.rc != 8u * (n - prefix - 4u)) { rc = -9; }
        }
    }
    if (lane == 0u) { status[si] = rc == 0 ? 0 : 1; }
}

// Persistent launch over the batch's rANS2 colour slices (longest first):
// wave k of the grid decodes entries k, k + W, ... (W waves in the grid).
// The CRC tables are shared by the workgroup's waves.
__global__ __launch_bounds__(256) void k_slice_decode(const TcgConsts* __restrict__ K,
                                                       const tcg_dplane* __restrict__ planes,
                                                       const tcg_dslice* __restrict__ slices,
                                                       int32_t* __restrict__ qbuf,
                                                       int32_t* __restrict__ status,
                                                       const tcg_work* __restrict__ work, uint32_t wave_words)
{
    __shared__ uint32_t sTab[256];
    __shared__ uint32_t sX2n[32];
    extern __shared__ uint32_t sWaves[];
    for (uint32_t i = threadIdx.x; i < 256u; i += blockDim.x) { sTab[i] = K->crc[i]; }
    if (threadIdx.x < 32u) { sX2n[threadIdx.x] = K->x2n[threadIdx.x]; }
    __syncthreads();
    const uint32_t lane = threadIdx.x & 63u;
    const WaveSlot ws = wave_slot();
    uint32_t* const mine = sWaves + ws.wave * wave_words;
    const uint32_t* const order = tcg_work_order(work);
    const uint32_t count = work->count;
    for (uint32_t i = ws.first; i < count; i += ws.stride) {
        slice_decode_one(K, planes, slices, qbuf, status, order[i], lane, sTab, sX2n, mine);
    }
}

// ---------------------------------------------------------------------------
// K6: dequantisation + inverse transform + level shift + clamp + clip-store.
// The entropy decoders store each block's coefficients column-major (natural
// index 8 r + c at 8 c + r). Tile = 32 blocks of one block row, 256 threads;
// each wave owns 8 blocks of the tile and a private LDS region, so no
// workgroup barrier is needed. Pass 1: lane (j, c) dequantises column c of
// block j and forms T[y][c] = sum_u M[u][y] G[u][c]; pass 2 (after a
// wave-level LDS transpose): lane (j, y) forms x[y][x] = sum_v T[y][v] M[v][x]
// and writes row y of the block as one 16-byte store. Every intermediate is
// an exact integer (f64 below 2^53 for bit depth <= 12, int64 above), so the
// order of the two passes does not change a sample. The dequantisation
// weights Q W 2^-32 come precomputed, transposed, with the quantiser context.
// ---------------------------------------------------------------------------
#define IQ_BLOCKS TCG_IQ_BLOCKS
#define IQ_WAVE_BLOCKS 8u
#define IQ_WAVES (IQ_BLOCKS / IQ_WAVE_BLOCKS)   // waves of a workgroup (one tile)

// 8-point inverse transform (exact even/odd form of out[x] = sum_v in[v] M[v][x])
__device__ static inline void idct8(const double* in, double* out) { tc_idct8_f64(in, out); }
__device__ static inline void idct8(const int64_t* in, int64_t* out) { tc_idct8_i64(in, out); }

// coefficients of column c of block j of a tile (thread t = 8 j + c: 32
// contiguous bytes), loaded ahead of the transform
struct IqRow { int4 r0, r1; };

__device__ static inline void iq_load(const tcg_dplane& P, uint32_t brow, uint32_t bx0, uint32_t nb,
                                      const int32_t* __restrict__ qbuf, IqRow& q)
{
    const uint32_t j = threadIdx.x >> 3, c = threadIdx.x & 7u;
    if (j < nb) {
        const int32_t* src = qbuf + ((size_t)P.blk0 + (size_t)brow * P.bcols + bx0 + j) * 64u + c * 8u;
        q.r0 = *(const int4*)src;
        q.r1 = *(const int4*)(src + 4);
    }
}

// Exact rounding of the f64 transform output by the reference's sign-split
// rule r = sign(v) floor((|v| + 2^31) / 2^32) for the integer v < 2^53 of the
// int64 transform: the f64 path carries t = v 2^-32 throughout (the weights
// are pre-scaled by 2^-32, a power of two, so every sum stays exact: the
// integers below 2^53 become multiples of 2^-32 below 2^21), t +
// copysign(1/2, t) is exact as well, and its truncation toward zero is
// sign(t) floor(|t| + 1/2).
__device__ static inline int32_t idct_round(double t)
{
    return (int32_t)(t + __builtin_copysign(0.5, t));
}

// x + mid clamped to [0, maxv] in one vector median
__device__ static inline int32_t idct_clamp(int32_t v, int32_t mid, int32_t maxv)
{
    int32_t r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(v + mid), "s"(maxv));
    return r;
}

__device__ static inline int32_t idct_round(int64_t v)
{
    return v >= 0 ? (int32_t)((v + ((int64_t)1 << 31)) >> 32) : -(int32_t)((-v + ((int64_t)1 << 31)) >> 32);
}

// LDS of one wave: the first pass's output of its 8 blocks (f64 or int64,
// 65-word pitch per block)
struct IqWaveLds {
    uint64_t B[IQ_WAVE_BLOCKS * 65];
};

// The 8 blocks bx0 .. bx0 + nb - 1 (nb <= 8) of block row brow, one wave.
// vec16: the plane's base and row pitch are 16-byte aligned.
template <bool WIDE>
__device__ static void dequant_idct_wave(const tcg_dplane& P, const tcg_qctx* __restrict__ qc,
                                         uint32_t brow, uint32_t bx0, uint32_t nb, bool conceal, bool vec16,
                                         const IqRow& pre, IqWaveLds& L, uint32_t lane)
{
    typedef typename std::conditional<WIDE, int64_t, double>::type acc_t;
    acc_t* sB = (acc_t*)L.B;
    const int32_t mid = qc->mid, maxv = qc->maxv;
    const int64_t fclamp = qc->fclamp;
    const uint32_t y0 = brow * 8u;
    if (conceal) {
        // fill the visible part of the wave's blocks with the neutral value
        const uint32_t x = bx0 * 8u + lane, xe = min(P.vis_w, (bx0 + nb) * 8u);
        uint16_t* out = (uint16_t*)P.dst;
        if (x < xe) {
            for (uint32_t y = y0; y < min(y0 + 8u, P.vis_h); ++y) { out[(size_t)y * P.stride + x] = (uint16_t)P.fill; }
        }
        return;
    }
    const uint32_t j = lane >> 3;
    {
        const uint32_t c = lane & 7u;   // pass 1: column c of block j
        const int4 r0 = pre.r0, r1 = pre.r1;
        int32_t qv[8] = { r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w };   // rows u = 0..7
        if (P.limit != 0u && P.limit < 63u) {
            // reduced decode: scan positions above the limit are dropped
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if ((uint32_t)kZigzagInv[u * 8 + c] > P.limit) { qv[u] = 0; }
            }
        }
        if (j < nb) {
            acc_t G[8];
            if (WIDE) {
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    int64_t d = (int64_t)qv[u] * (int64_t)qc->Q[u * 8 + c];
                    d = d > fclamp ? fclamp : (d < -fclamp ? -fclamp : d);
                    d = d > (1 << 30) ? (1 << 30) : (d < -(1 << 30) ? -(1 << 30) : d);
                    G[u] = (acc_t)d * (acc_t)kTW[u * 8 + c];
                }
            } else {
                // lanes holding a coefficient that may reach the clamp: |q|
                // above the smallest per-position limit floor(fclamp / Q)
                // (accepted levels are below 2^27, so |q| does not wrap)
                uint32_t amax = 0u;
#pragma unroll
                for (int u = 0; u < 8; ++u) { amax = max(amax, (uint32_t)abs(qv[u])); }
                if (__builtin_expect(__ballot(amax > qc->qlim_min) == 0ull, 1)) {
                    // |q| <= floor(fclamp / Q) keeps |q Q| within the clamp,
                    // so the coefficient dequantises as q (Q W 2^-32) exactly
                    const double2* qw = (const double2*)(qc->qwt + c * 8u);
#pragma unroll
                    for (int u = 0; u < 8; u += 2) {
                        const double2 w2 = qw[u / 2];
                        G[u] = (acc_t)((double)qv[u] * w2.x);
                        G[u + 1] = (acc_t)((double)qv[u + 1] * w2.y);
                    }
                } else {
                    // |q| < 2^27 and Q < 2^26, so q Q is an integer below 2^53
                    // and exact in f64; the clamp bounds it by 2^25 and the
                    // weight by 2^14, so the product with the (2^-32-scaled)
                    // weight is exact as well
                    const double fc = (double)fclamp;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        double d = (double)qv[u] * (double)qc->Q[u * 8 + c];
                        d = fmin(fmax(d, -fc), fc);
                        G[u] = (acc_t)(d * ((double)kTW[u * 8 + c] * 0x1p-32));
                    }
                }
            }
            acc_t Tc[8];
            idct8(G, Tc);
#pragma unroll
            for (int y = 0; y < 8; ++y) { sB[j * 65 + y * 8 + c] = Tc[y]; }
        }
    }
    wave_sync();   // the wave's own LDS region: its lanes exchange columns for rows
    if (j < nb) {
        const uint32_t y = lane & 7u;   // pass 2: row y of block j
        acc_t Tr[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) { Tr[v] = sB[j * 65 + y * 8 + v]; }
        acc_t Xv[8];
        idct8(Tr, Xv);
        uint32_t px[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            px[k] = (uint32_t)idct_clamp(idct_round(Xv[2 * k]), mid, maxv) |
                    (uint32_t)idct_clamp(idct_round(Xv[2 * k + 1]), mid, maxv) << 16;
        }
        const uint32_t x0 = (bx0 + j) * 8u;
        if (y0 + y < P.vis_h) {
            __attribute__((address_space(1))) uint16_t* dst =
                (__attribute__((address_space(1))) uint16_t*)(uintptr_t)P.dst + (size_t)(y0 + y) * P.stride + x0;
            if (vec16 && x0 + 8u <= P.vis_w) {
                typedef uint32_t iq_u32x4 __attribute__((ext_vector_type(4)));
                *(__attribute__((address_space(1))) iq_u32x4*)dst = iq_u32x4{ px[0], px[1], px[2], px[3] };
            } else {
#pragma unroll
                for (int x = 0; x < 8; ++x) {
                    if (x0 + (uint32_t)x < P.vis_w) { dst[x] = (uint16_t)(px[x >> 1] >> (16 * (x & 1))); }
                }
            }
        }
    }
    wave_sync();   // pass-2 reads done before the region is written again
}

// grid: x = group of TPW tiles (32 blocks of a block row each) within the
// plane, y = plane of the batch. Every tile's row record (quantiser context,
// slice) and coefficients are loaded first (TPW x 8 KiB in flight per
// workgroup); then each wave transforms its 8 blocks of each tile on its own.
// At most 80 VGPRs (six waves per SIMD, no spills): the wave count, not the
// registers the compiler would otherwise keep live, sets how much of the HBM
// latency the kernel hides.
template <uint32_t TPW>
__global__ __launch_bounds__(64u * IQ_WAVES) __attribute__((amdgpu_waves_per_eu(6))) void k_dequant_idct(const tcg_qctx* __restrict__ qctxs,
                                                       const tcg_dplane* __restrict__ planes,
                                                       const tcg_rowrec* __restrict__ rowrec,
                                                       uint32_t max_rows,
                                                       const int32_t* __restrict__ status,
                                                       const int32_t* __restrict__ qbuf)
{
    __shared__ IqWaveLds L[IQ_WAVES];
    const uint32_t pi = blockIdx.y;
    const tcg_dplane P = planes[pi];
    if (P.kind != 0u) { return; }
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t tiles_per_row = (P.bcols + IQ_BLOCKS - 1u) / IQ_BLOCKS;
    const uint32_t ntiles = P.brows * tiles_per_row;
    if (blockIdx.x * TPW >= ntiles) { return; }
    const bool vec16 = ((P.dst | ((uint64_t)P.stride * 2u)) & 15u) == 0u;
    uint32_t brow[TPW], bx0[TPW];
    tcg_rowrec rr[TPW];
#pragma unroll
    for (uint32_t k = 0; k < TPW; ++k) {
        const uint32_t local = min(blockIdx.x * TPW + k, ntiles - 1u);
        brow[k] = tc_qdiv32(local, P.tile_m, P.tile_l);
        bx0[k] = (local - brow[k] * tiles_per_row) * IQ_BLOCKS;
        rr[k] = rowrec[(size_t)pi * max_rows + brow[k]];
    }
    IqRow pre[TPW];
#pragma unroll
    for (uint32_t k = 0; k < TPW; ++k) {
        if (blockIdx.x * TPW + k < ntiles) {
            iq_load(P, brow[k], bx0[k], min((uint32_t)IQ_BLOCKS, P.bcols - bx0[k]), qbuf, pre[k]);
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < TPW; ++k) {
        if (blockIdx.x * TPW + k >= ntiles) { break; }
        const uint32_t wb0 = bx0[k] + wave * IQ_WAVE_BLOCKS;   // the wave's first block
        if (wb0 >= P.bcols) { continue; }
        const uint32_t nb = min(IQ_WAVE_BLOCKS, P.bcols - wb0);
        const tcg_qctx* qc = &qctxs[rr[k].qctx];
        const bool conceal = status[rr[k].slice] != 0;
        if (qc->bit_depth > 12u) {
            dequant_idct_wave<true>(P, qc, brow[k], wb0, nb, conceal, vec16, pre[k], L[wave], lane);
        } else {
            dequant_idct_wave<false>(P, qc, brow[k], wb0, nb, conceal, vec16, pre[k], L[wave], lane);
        }
    }
}

// ---------------------------------------------------------------------------
// K7: output deblocking, in place. Vertical edges sit at x0 = g k (g = 8, or 4
// on an exact 2:1 target), horizontal edges at y0 = g r, and an edge reads and
// writes the two samples on each side of it. The plane therefore splits into
// g x g tiles offset by g/2 in both directions, (x0 - g/2 .. x0 + g/2) x
// (y0 - g/2 .. y0 + g/2), each holding exactly one vertical edge segment (g
// rows) and one horizontal edge segment (g columns) and every sample those
// segments read: filtering the vertical segment of a tile and then its
// horizontal segment gives the result of the whole-plane vertical pass
// followed by the whole-plane horizontal pass (deblock.c:163-212), and tiles
// share no sample, so they run independently in place. One thread per tile,
// one wave per tile row, taking the row in runs of consecutive tiles. With g = 8 on a plane
// whose base and row pitch are 16-byte aligned the wave reads and writes each
// row as aligned 16-byte chunks and passes half-tiles between neighbouring
// lanes (deblock_chunks); otherwise each thread reads its tile's half-rows
// with 8-byte (g = 8) or 4-byte (g = 4) vector loads where the alignment
// allows, or sample by sample (deblock_tile).
// ---------------------------------------------------------------------------
__device__ static inline uint32_t db_absu(int32_t v) { return (uint32_t)(v < 0 ? -v : v); }
__device__ static inline int32_t db_clampv(int32_t v, int32_t m) { return v < 0 ? 0 : (v > m ? m : v); }

__device__ static inline void db_edge(int32_t& p1, int32_t& p0, int32_t& q0, int32_t& q1, int32_t alpha,
                                      int32_t beta, int32_t clip, int32_t maxv)
{
    if (!(db_absu(q0 - p0) < (uint32_t)alpha && db_absu(p1 - p0) < (uint32_t)beta &&
          db_absu(q1 - q0) < (uint32_t)beta)) {
        return;
    }
    // arithmetic shifts of negative values, as the reference's gcc build
    int32_t d = (12 * (q0 - p0) + 4 * (p1 - q1) + 8) >> 4;
    const int32_t delta = q0 - p0;
    const int32_t half = delta >= 0 ? (delta >> 1) : -((-delta) >> 1);
    if (half >= 0) { d = d > half ? half : (d < 0 ? 0 : d); }
    else { d = d < half ? half : (d > 0 ? 0 : d); }
    d = d < -clip ? -clip : (d > clip ? clip : d);
    const int32_t ds = d >> 1;
    const int32_t np0 = db_clampv(p0 + d, maxv), nq0 = db_clampv(q0 - d, maxv);
    const int32_t np1 = db_clampv(p1 + ds, maxv), nq1 = db_clampv(q1 - ds, maxv);
    p0 = np0; q0 = nq0; p1 = np1; q1 = nq1;
}

#define DB_TX TCG_DB_TX         // lanes of a wave (tiles of one tile row)
#define DB_WAVES TCG_DB_WAVES   // waves of a workgroup, along one tile row

typedef __attribute__((address_space(1))) uint8_t gbyte;

// half-row of a tile (g/2 samples) as packed sample pairs, at a byte offset
// of the plane (offsets are 32-bit, so the address is the plane base in
// scalar registers plus a 32-bit vector offset)
template <uint32_t G, bool VEC>
__device__ static inline uint2 db_load_half(const gbyte* base, uint32_t off)
{
    const __attribute__((address_space(1))) uint16_t* p = (const __attribute__((address_space(1))) uint16_t*)(base + off);
    if (G == 8u) {
        if (VEC) {
            const uint64_t q = *(const __attribute__((address_space(1))) uint64_t*)(base + off);
            return make_uint2((uint32_t)q, (uint32_t)(q >> 32));
        }
        return make_uint2(p[0] | (uint32_t)p[1] << 16, p[2] | (uint32_t)p[3] << 16);
    }
    if (VEC) { return make_uint2(*(const __attribute__((address_space(1))) uint32_t*)(base + off), 0u); }
    return make_uint2(p[0] | (uint32_t)p[1] << 16, 0u);
}

template <uint32_t G, bool VEC>
__device__ static inline void db_store_half(gbyte* base, uint32_t off, const int32_t* in)
{
    if (G == 8u) {
        const uint2 d = make_uint2((uint32_t)in[0] | (uint32_t)in[1] << 16, (uint32_t)in[2] | (uint32_t)in[3] << 16);
        if (VEC) { *(__attribute__((address_space(1))) uint64_t*)(base + off) = (uint64_t)d.y << 32 | d.x; }
        else {
            __attribute__((address_space(1))) uint16_t* p = (__attribute__((address_space(1))) uint16_t*)(base + off);
            p[0] = (uint16_t)d.x; p[1] = (uint16_t)(d.x >> 16); p[2] = (uint16_t)d.y; p[3] = (uint16_t)(d.y >> 16);
        }
    } else {
        const uint32_t d = (uint32_t)in[0] | (uint32_t)in[1] << 16;
        if (VEC) { *(__attribute__((address_space(1))) uint32_t*)(base + off) = d; }
        else {
            __attribute__((address_space(1))) uint16_t* p = (__attribute__((address_space(1))) uint16_t*)(base + off);
            p[0] = (uint16_t)d; p[1] = (uint16_t)(d >> 16);
        }
    }
}

// One corner-centred tile (k, r): half-rows outside the plane read as zeros
// from a clamped in-plane address (no branch around the loads) and are not
// written back. Byte offsets inside a plane fit 32 bits (planes are below
// 2^31 samples).
template <uint32_t G, bool VEC>
__device__ static inline void deblock_tile(const tcg_dplane& P, int32_t k, int32_t r)
{
    constexpr int32_t HG = (int32_t)G / 2;
    const int32_t w = (int32_t)P.vis_w, h = (int32_t)P.vis_h;
    const int32_t x0 = k * (int32_t)G - HG, y0 = r * (int32_t)G - HG;  // tile origin
    gbyte* const base = (gbyte*)(uintptr_t)P.dst;
    const uint32_t pitch = P.stride * 2u;
    bool hin[2];
    uint32_t xb[2];
#pragma unroll
    for (int32_t half = 0; half < 2; ++half) {
        const int32_t xs = x0 + half * HG;
        hin[half] = xs >= 0 && xs < w;   // a half-row is all in or all out
        xb[half] = hin[half] ? (uint32_t)xs * 2u : 0u;
    }
    // every load first (unconditional, clamped addresses), then the masking
    uint2 raw[G][2];
#pragma unroll
    for (int32_t i = 0; i < (int32_t)G; ++i) {
        const int32_t y = y0 + i;
        const uint32_t ro = (y >= 0 && y < h) ? (uint32_t)y * pitch : 0u;
#pragma unroll
        for (int32_t half = 0; half < 2; ++half) { raw[i][half] = db_load_half<G, VEC>(base, ro + xb[half]); }
    }
    int32_t v[G][G];
#pragma unroll
    for (int32_t i = 0; i < (int32_t)G; ++i) {
        const bool rin = y0 + i >= 0 && y0 + i < h;
#pragma unroll
        for (int32_t half = 0; half < 2; ++half) {
            const bool in = rin && hin[half];
            const uint32_t a = in ? raw[i][half].x : 0u, b = in ? raw[i][half].y : 0u;
            v[i][half * HG + 0] = (int32_t)(a & 0xFFFFu);
            v[i][half * HG + 1] = (int32_t)(a >> 16);
            if (G == 8u) {
                v[i][half * HG + 2] = (int32_t)(b & 0xFFFFu);
                v[i][half * HG + 3] = (int32_t)(b >> 16);
            }
        }
    }
    const int32_t alpha = (int32_t)P.db_alpha, beta = alpha / 2 + 2, clip = alpha >> 1;
    const int32_t maxv = (int32_t)P.maxv;
    // vertical edge x = g k (g <= x, x + 1 < w), every row of the tile
    if (k >= 1 && (k + 1) * (int32_t)G <= w) {
#pragma unroll
        for (int32_t i = 0; i < (int32_t)G; ++i) {
            if (y0 + i >= 0 && y0 + i < h) {
                db_edge(v[i][HG - 2], v[i][HG - 1], v[i][HG], v[i][HG + 1], alpha, beta, clip, maxv);
            }
        }
    }
    // horizontal edge y = g r (1 <= r <= h/g - 1), every column of the tile
    if (r >= 1 && (r + 1) * (int32_t)G <= h) {
#pragma unroll
        for (int32_t j = 0; j < (int32_t)G; ++j) {
            if (x0 + j >= 0 && x0 + j < w) {
                db_edge(v[HG - 2][j], v[HG - 1][j], v[HG][j], v[HG + 1][j], alpha, beta, clip, maxv);
            }
        }
    }
#pragma unroll
    for (int32_t i = 0; i < (int32_t)G; ++i) {
        const int32_t y = y0 + i;
        if (y < 0 || y >= h) { continue; }
        const uint32_t ro = (uint32_t)y * pitch;
#pragma unroll
        for (int32_t half = 0; half < 2; ++half) {
            if (hin[half]) { db_store_half<G, VEC>(base, ro + xb[half], &v[i][half * HG]); }
        }
    }
}

typedef uint32_t db_u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t db_u32x2 __attribute__((ext_vector_type(2)));

// a lane's value as seen by the lane above it (lane 0 reads 0) and by the
// lane below it (lane 63 reads 0): DPP whole-wave shifts by one lane
__device__ static inline uint32_t db_from_below(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138 /* wave_shr:1 */, 0xf, 0xf, false);
}
__device__ static inline uint32_t db_from_above(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130 /* wave_shl:1 */, 0xf, 0xf, false);
}

// g = 8, plane base and row pitch 16-byte aligned: the wave covers tiles
// t0 .. t0 + 62 of tile row r. Lane l reads the aligned chunk c = t0 + l - 1
// of each of the tile row's 8 rows (samples 8c .. 8c + 7: the right half of
// tile c and the left half of tile c + 1), so each row load of the wave is
// one contiguous 1 KiB span. Lane l >= 1 owns tile c: its left half arrives
// from lane l - 1 and, once filtered, goes back to lane l - 1, which writes
// the whole chunk. Lane 0 only supplies and writes back tile t0's left half,
// and lane 63 writes only its own tile's right half (the upper half of its
// chunk is the next wave's first tile). Rows and chunks outside the plane are
// read from a clamped in-plane address and reach no filtered or stored
// sample: no edge is filtered there and no such chunk is written. Called
// with every lane of the wave active (the DPP shifts read all 64 lanes).
__device__ static inline void deblock_chunks(const tcg_dplane& P, int32_t t0, int32_t r, int32_t lane)
{
    constexpr int32_t G = 8, HG = 4;
    const int32_t w = (int32_t)P.vis_w, h = (int32_t)P.vis_h;
    const int32_t c = t0 + lane - 1, x0 = c * G - HG, y0 = r * G - HG;
    gbyte* const base = (gbyte*)(uintptr_t)P.dst;
    const uint32_t pitch = P.stride * 2u;
    const bool cin = c >= 0 && c < w / G;
    const uint32_t cb = cin ? (uint32_t)c * 16u : 0u;
    db_u32x4 raw[G];
#pragma unroll
    for (int32_t i = 0; i < G; ++i) {
        const int32_t y = y0 + i;
        const uint32_t ro = (y >= 0 && y < h) ? (uint32_t)y * pitch : 0u;
        raw[i] = *(const __attribute__((address_space(1))) db_u32x4*)(base + ro + cb);
    }
    int32_t v[G][G];
#pragma unroll
    for (int32_t i = 0; i < G; ++i) {
        const uint32_t a = db_from_below(raw[i].z), b = db_from_below(raw[i].w);
        v[i][0] = (int32_t)(a & 0xFFFFu);
        v[i][1] = (int32_t)(a >> 16);
        v[i][2] = (int32_t)(b & 0xFFFFu);
        v[i][3] = (int32_t)(b >> 16);
        v[i][4] = (int32_t)(raw[i].x & 0xFFFFu);
        v[i][5] = (int32_t)(raw[i].x >> 16);
        v[i][6] = (int32_t)(raw[i].y & 0xFFFFu);
        v[i][7] = (int32_t)(raw[i].y >> 16);
    }
    const int32_t alpha = (int32_t)P.db_alpha, beta = alpha / 2 + 2, clip = alpha >> 1;
    const int32_t maxv = (int32_t)P.maxv;
    if (c >= 1 && (c + 1) * G <= w) {
#pragma unroll
        for (int32_t i = 0; i < G; ++i) {
            if (y0 + i >= 0 && y0 + i < h) {
                db_edge(v[i][HG - 2], v[i][HG - 1], v[i][HG], v[i][HG + 1], alpha, beta, clip, maxv);
            }
        }
    }
    if (r >= 1 && (r + 1) * G <= h) {
#pragma unroll
        for (int32_t j = 0; j < G; ++j) {
            if (x0 + j >= 0 && x0 + j < w) {
                db_edge(v[HG - 2][j], v[HG - 1][j], v[HG][j], v[HG + 1][j], alpha, beta, clip, maxv);
            }
        }
    }
#pragma unroll
    for (int32_t i = 0; i < G; ++i) {
        const int32_t y = y0 + i;
        const uint32_t lo0 = (uint32_t)v[i][4] | (uint32_t)v[i][5] << 16;
        const uint32_t lo1 = (uint32_t)v[i][6] | (uint32_t)v[i][7] << 16;
        const uint32_t hi0 = db_from_above((uint32_t)v[i][0] | (uint32_t)v[i][1] << 16);
        const uint32_t hi1 = db_from_above((uint32_t)v[i][2] | (uint32_t)v[i][3] << 16);
        if (y < 0 || y >= h || !cin) { continue; }
        gbyte* const row = base + (uint32_t)y * pitch + cb;
        if (lane >= 1 && lane <= 62) { *(__attribute__((address_space(1))) db_u32x4*)row = db_u32x4{lo0, lo1, hi0, hi1}; }
        else if (lane == 0) { *(__attribute__((address_space(1))) db_u32x2*)(row + 8) = db_u32x2{hi0, hi1}; }
        else { *(__attribute__((address_space(1))) db_u32x2*)row = db_u32x2{lo0, lo1}; }
    }
}

// grid: y = tile row, z = plane of the batch; a workgroup of DB_WAVES waves
// per tile row, wave w taking runs w, w + DB_WAVES, ... of DB_TX - 1 tiles
// along the row (DB_TX on the per-thread tile path). Neighbouring runs share
// the cache line at their boundary, which stays in the workgroup's L2.
template <uint32_t G>
__global__ __launch_bounds__(DB_TX * DB_WAVES) void k_deblock(const tcg_dplane* __restrict__ planes)
{
    const tcg_dplane P = planes[blockIdx.z];
    if (!P.deblock || (P.db_unit ? P.db_unit : 8u) != G) { return; }
    const int32_t lane = (int32_t)(threadIdx.x % DB_TX), wave = (int32_t)(threadIdx.x / DB_TX);
    const int32_t r = (int32_t)blockIdx.y;
    if (r > (int32_t)P.vis_h / (int32_t)G) { return; }
    const int32_t last = (int32_t)P.vis_w / (int32_t)G;   // tiles 0 .. last of the row
    const uint64_t align = P.dst | ((uint64_t)P.stride * 2u);
    if (G == 8u && (align & 15u) == 0u) {
        for (int32_t t0 = wave * (DB_TX - 1); t0 <= last; t0 += DB_WAVES * (DB_TX - 1)) {
            deblock_chunks(P, t0, r, lane);
        }
        return;
    }
    // vector access needs the tile's half-rows (g/2 samples) aligned: the
    // plane's base and pitch at g bytes (w and h are multiples of g)
    const bool vec = (align & (uint64_t)(G - 1u)) == 0u;
    for (int32_t k0 = wave * DB_TX; k0 <= last; k0 += DB_WAVES * DB_TX) {
        const int32_t k = k0 + lane;
        if (k > last) { break; }
        if (vec) { deblock_tile<G, true>(P, k, r); }
        else { deblock_tile<G, false>(P, k, r); }
    }
}

// Target-size decode: target sample (tx, ty) of a plane takes the source
// sample (floor(tx vw / dw), floor(ty vh / dh)) of the reconstructed
// source-size plane (deblocked in place afterwards when the target filters).
__global__ __launch_bounds__(256) void k_sample(const tcg_dplane* __restrict__ srcs,
                                                const tcg_dplane* __restrict__ dsts)
{
    const uint32_t pi = blockIdx.y;
    const tcg_dplane D = dsts[pi];
    if (!D.sample) { return; }
    const tcg_dplane S = srcs[pi];
    const uint16_t* src = (const uint16_t*)S.dst;
    uint16_t* out = (uint16_t*)D.dst;
    const uint64_t n = (uint64_t)D.vis_w * D.vis_h;
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) {
        const uint32_t ty = (uint32_t)(i / D.vis_w), tx = (uint32_t)(i - (uint64_t)ty * D.vis_w);
        const uint32_t sy = (uint32_t)(((uint64_t)ty * S.vis_h) / D.vis_h);
        const uint32_t sx = (uint32_t)(((uint64_t)tx * S.vis_w) / D.vis_w);
        out[(size_t)ty * D.stride + tx] = src[(size_t)sy * S.stride + sx];
    }
}

// symbol pairs of the batch's rANS2 colour slices (timed sequences only)
__global__ __launch_bounds__(256) void k_count_pairs_dec(const tcg_dplane* __restrict__ planes,
                                                         const tcg_dslice* __restrict__ slices,
                                                         const int32_t* __restrict__ q, unsigned long long* counter)
{
    const tcg_dslice S = slices[blockIdx.x];
    if (S.kind != 0u || S.mode != TCG_MODE_RANS2) { return; }
    tc_count_pairs(q, S.blk0, S.h * planes[S.plane].bcols, counter);
}

// ---------------------------------------------------------------------------
// host launch
// ---------------------------------------------------------------------------
extern "C" int32_t tcg_decode(tcg_ctx* c, const tcg_qctx* qctx_dev, uint32_t nplanes, const tcg_dplane* planes_dev,
                              uint32_t nslices, const tcg_dslice* slices_dev, const tcg_rowrec* rowrec_dev,
                              uint32_t max_rows, uint64_t qbuf, uint32_t max_plane_tiles,
                              int32_t* slice_status_dev, uint32_t max_db_tiles, const tcg_batch_info* info)
{
    hipStream_t st = c->stream;
    const TcgConsts* K = tcg_consts_dev();
    c->kind = 2;
    const uint32_t wpg = info->ent_wpg ? info->ent_wpg : 1u;
    const uint32_t iq_tpw = info->tile_tpw ? info->tile_tpw : 1u;
    const uint32_t wave_words = dec_wave_words(info->max_cols);
    const uint32_t groups = info->ent_groups ? info->ent_groups : (info->ent_count + wpg - 1u) / wpg;
    auto launch = [&]() {
        tcg_mark(c, 0);
        if ((info->kinds & TCG_HAS_RANS2) && groups > 0u) {
            hipLaunchKernelGGL(k_slice_decode, dim3(groups), dim3(64u * wpg), (size_t)wave_words * 4u * wpg, st, K,
                               planes_dev, slices_dev, (int32_t*)qbuf, slice_status_dev, info->work, wave_words);
        }
        if (info->kinds & TCG_HAS_BITS) {
            tcg_bits_decode_launch(st, K, planes_dev, slices_dev, nslices, (int32_t*)qbuf, slice_status_dev,
                                   info->max_cols);
        }
        if (info->kinds & TCG_HAS_ALPHA) {
            tcg_alpha_decode_launch(st, planes_dev, slices_dev, nslices, (int32_t*)qbuf, slice_status_dev);
        }
        tcg_mark(c, 1);
        if (max_plane_tiles > 0u) {
            if (iq_tpw == 4u) {
                hipLaunchKernelGGL(k_dequant_idct<4>, dim3((max_plane_tiles + 3u) / 4u, nplanes), dim3(64u * IQ_WAVES), 0, st,
                                   qctx_dev, planes_dev, rowrec_dev, max_rows, slice_status_dev, (const int32_t*)qbuf);
            } else if (iq_tpw == 2u) {
                hipLaunchKernelGGL(k_dequant_idct<2>, dim3((max_plane_tiles + 1u) / 2u, nplanes), dim3(64u * IQ_WAVES), 0, st,
                                   qctx_dev, planes_dev, rowrec_dev, max_rows, slice_status_dev, (const int32_t*)qbuf);
            } else {
                hipLaunchKernelGGL(k_dequant_idct<1>, dim3(max_plane_tiles, nplanes), dim3(64u * IQ_WAVES), 0, st, qctx_dev,
                                   planes_dev, rowrec_dev, max_rows, slice_status_dev, (const int32_t*)qbuf);
            }
        }
        tcg_mark(c, 2);
        if (max_db_tiles > 0u) {
            hipLaunchKernelGGL(k_deblock<8>, dim3(1, max_db_tiles & 0xFFFFu, nplanes), dim3(DB_TX * DB_WAVES), 0,
                               st, planes_dev);
        }
        tcg_mark(c, 3);
        unsigned long long* pairs = (info->kinds & TCG_HAS_RANS2) ? tcg_pairs_counter(c) : nullptr;
        if (pairs != nullptr) {
            hipLaunchKernelGGL(k_count_pairs_dec, dim3(nslices), dim3(256), 0, st, planes_dev, slices_dev,
                               (const int32_t*)qbuf, pairs);
        }
    };
    const uint64_t key[TCG_GRAPH_KEY] = {
        2u, (uint64_t)(uintptr_t)qctx_dev, nplanes, (uint64_t)(uintptr_t)planes_dev, nslices,
        (uint64_t)(uintptr_t)slices_dev, (uint64_t)(uintptr_t)rowrec_dev, max_rows, qbuf, max_plane_tiles,
        (uint64_t)(uintptr_t)slice_status_dev, max_db_tiles, info->max_cols, info->kinds, (uint64_t)(uintptr_t)K,
        (uint64_t)wpg << 32 | groups, (uint64_t)(uintptr_t)info->work, iq_tpw, 0u, 0u };
    return tcg_run(c, key, launch);
}

uint32_t tcg_entropy_capacity_dec(uint32_t wpg, uint32_t max_cols)
{
    const uint32_t wave_words = dec_wave_words(max_cols);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_slice_decode, (int)(64u * wpg),
                                                     (size_t)wave_words * 4u * wpg) != hipSuccess) {
        (void)hipGetLastError();
        return 0u;
    }
    return (uint32_t)per_cu * tcg_cu_count();
}

extern "C" int32_t tcg_decode_sample(tcg_ctx* c, const tcg_dplane* src_dev, const tcg_dplane* dst_dev,
                                     uint32_t nplanes, uint32_t max_px, uint32_t max_db_tiles)
{
    if (nplanes == 0u || max_px == 0u) { return 0; }
    const uint32_t gx = min((max_px + 255u) / 256u, 8192u);
    hipLaunchKernelGGL(k_sample, dim3(gx, nplanes), dim3(256), 0, c->stream, src_dev, dst_dev);
    if (max_db_tiles > 0u) {
        hipLaunchKernelGGL(k_deblock<4>, dim3(1, max_db_tiles & 0xFFFFu, nplanes), dim3(DB_TX * DB_WAVES), 0,
                           c->stream, dst_dev);
    }
    return tcg_check("target-size sample launch");
}
