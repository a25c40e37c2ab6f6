// gfx950 encode kernels: fused level-shift + 8x8 integer DCT + dead-zone
// quantisation, per-slice rANS2 ("V7-R2") entropy coding with one wavefront
// per slice, packet offset scan, and payload copy + CRC32 + slice headers.
//
// Reference semantics (paths relative to the reference tree):
//   transform   native/src/transform/transform.c:18-41, simd/dispatch.c:402-413
//   quantiser   native/src/transform/quant.c:310-328
//   tokens      native/src/codec/codec.c:1702-1848
//   rANS2       native/src/codec/codec.c:2122-2515, native/src/entropy/rans.c
//   framing     native/src/bitstream/frame_header.c:590-631, slice_map.c:9-21
#include <hip/hip_runtime.h>

#include "tcg.h"
#include "tc_math.h"
#include "tc_dev.h"

// ---------------------------------------------------------------------------
// K1: fused FDCT + quantisation.
// One workgroup = 8 x FQ_BLOCKS threads = one tile of FQ_BLOCKS blocks along a
// block row. Pass 1 (rows): thread (r, j) loads row r of block j (8 u16 =
// 16 B; the threads of a row read 16 x FQ_BLOCKS contiguous bytes) and
// computes A[r][v]. Pass 2 (columns) through LDS: thread (v, j) computes
// F[u][v], u = 0..7, and quantises; the coefficients are staged in LDS in
// scan (zigzag) order and leave as one contiguous span of 256 B per block.
// ---------------------------------------------------------------------------
#define FQ_BLOCKS TCG_FQ_BLOCKS
#define FQ_THREADS (8u * FQ_BLOCKS)
#ifndef FQ_NT_STORE
#define FQ_NT_STORE 0
#endif
// LDS pitch of one block's 64 words: 65 puts the blocks of a half-wave in
// different banks for the row/column transposes
#define FQ_PITCH 65u

struct FqTile {
    uint32_t brow, bx0, nb;
    bool live;
};

__device__ static inline FqTile fq_tile(const tcg_eplane& P, uint32_t local)
{
    FqTile T;
    const uint32_t tiles_per_row = (P.bcols + FQ_BLOCKS - 1u) / FQ_BLOCKS;
    T.live = local < P.brows * tiles_per_row;
    const uint32_t brow = tc_qdiv32(local, P.tile_m, P.tile_l);   // local / tiles_per_row
    T.brow = T.live ? brow : 0u;
    T.bx0 = T.live ? (local - brow * tiles_per_row) * FQ_BLOCKS : 0u;
    T.nb = T.live ? min((uint32_t)FQ_BLOCKS, P.bcols - T.bx0) : 0u;
    return T;
}

// level-shifted samples of row r of block j of the tile (rows past the
// visible height repeat the last row, columns past the width the last column)
__device__ static inline void fq_load(const tcg_eplane& P, const FqTile& T, int32_t mid, uint32_t t,
                                      int32_t* x)
{
    const uint32_t r = t / FQ_BLOCKS, j = t % FQ_BLOCKS;
    if (!T.live || j >= T.nb) { return; }
    const uint16_t* src = (const uint16_t*)P.src;
    uint32_t y = T.brow * 8u + r;
    if (y >= P.vis_h) { y = P.vis_h - 1u; }
    const uint32_t x0 = (T.bx0 + j) * 8u;
    const uint16_t* row = src + (size_t)y * P.stride;
    if (x0 + 8u <= P.vis_w && ((((uintptr_t)(row + x0)) & 15u) == 0u)) {
        const uint4 v = *(const uint4*)(row + x0);
        const uint32_t w[4] = { v.x, v.y, v.z, v.w };
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            x[2 * k] = (int32_t)(w[k] & 0xFFFFu) - mid;
            x[2 * k + 1] = (int32_t)(w[k] >> 16) - mid;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            uint32_t xx = x0 + (uint32_t)k;
            if (xx >= P.vis_w) { xx = P.vis_w - 1u; }
            x[k] = (int32_t)row[xx] - mid;
        }
    }
}

// The 8-point forward transform of tc_fdct8_i32 with 24-bit multiplies
// (v_mul_i32_i24, full rate, where a 32-bit multiply issues at quarter rate).
// Exact when every multiplicand is below 2^23 in magnitude: for samples of at
// most 12 bits (level-shifted, |x| <= 2048) the row pass's sums and
// differences stay below 2^15 and its outputs below 13 * 2^14 < 2^18; the
// column pass's sums of eight such values stay below 2^21, and every product
// (|13 * sum| < 2^25) fits the 32 bits the instruction returns.
__device__ static inline void fdct8_i24(const int32_t* in, int32_t* out)
{
    const int32_t s0 = in[0] + in[7], s1 = in[1] + in[6], s2 = in[2] + in[5], s3 = in[3] + in[4];
    const int32_t d0 = in[0] - in[7], d1 = in[1] - in[6], d2 = in[2] - in[5], d3 = in[3] - in[4];
    const int32_t e0 = s0 + s3, e1 = s1 + s2, f0 = s0 - s3, f1 = s1 - s2;
    out[0] = __mul24(13, e0 + e1);
    out[4] = __mul24(9, e0 - e1);
    out[2] = __mul24(12, f0) + __mul24(5, f1);
    out[6] = __mul24(5, f0) - __mul24(12, f1);
    out[1] = __mul24(13, d0) + __mul24(11, d1) + __mul24(7, d2) + d3;
    out[3] = __mul24(11, d0) - __mul24(4, d1) - __mul24(13, d2) - __mul24(8, d3);
    out[5] = __mul24(8, d0) - __mul24(13, d1) + __mul24(4, d2) + __mul24(11, d3);
    out[7] = d0 - __mul24(7, d1) + __mul24(11, d2) - __mul24(13, d3);
}

template <bool WIDE>
__device__ static inline void fdct8_pass(const int32_t* in, int32_t* out)
{
    if (WIDE) { tc_fdct8_i32(in, out); }
    else { fdct8_i24(in, out); }
}

template <bool WIDE>
__device__ static void fdct_quant_tile(const tcg_eplane& P, const FqTile& T, int32_t fclamp, int32_t* x,
                                       int32_t* __restrict__ qbuf, int32_t* sA, const uint32_t* sQm,
                                       const uint32_t* sQl, const uint32_t* sBias, const uint8_t* sZzc)
{
    const uint32_t t = threadIdx.x;
    const uint32_t nb = T.nb;
    // ---- pass 1: rows ----
    {
        const uint32_t r = t / FQ_BLOCKS, j = t % FQ_BLOCKS;
        if (j < nb) {
            if (!WIDE) {
#pragma unroll
                for (int k = 0; k < 8; ++k) { x[k] = (int32_t)(int16_t)x[k]; }
            }
            int32_t A[8];
            fdct8_pass<WIDE>(x, A);
#pragma unroll
            for (int v = 0; v < 8; ++v) { sA[j * FQ_PITCH + r * 8 + v] = A[v]; }
        }
    }
    __syncthreads();
    // ---- pass 2: columns + quantisation ----
    int32_t qv[8];
    const uint32_t v = t / FQ_BLOCKS, j = t % FQ_BLOCKS;
    if (j < nb) {
        int32_t a[8];
#pragma unroll
        for (int y = 0; y < 8; ++y) { a[y] = sA[j * FQ_PITCH + y * 8 + v]; }
        int32_t Fc[8];
        fdct8_pass<WIDE>(a, Fc);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int32_t F = Fc[u];
            const uint32_t nat = (uint32_t)u * 8u + v;
            if (WIDE) { F = predistort(F, nat); }
            const int32_t c = min(max(F, -fclamp), fclamp);   // one v_med3_i32
            const uint32_t mag = (uint32_t)(c < 0 ? -c : c);
            // Q/2 - dz >= 0 (dz = Q/4 for AC, 0 for DC), so |F| + Q/2 - dz never
            // goes below zero and needs no saturation
            const uint32_t ne = mag + sBias[nat];
            const uint32_t qm = tc_qdiv32(ne, sQm[nat], sQl[nat]);
            qv[u] = c < 0 ? -(int32_t)qm : (int32_t)qm;
        }
    }
    __syncthreads();
    if (j < nb) {
        // scan positions of this column's eight coefficients (one 8-byte read)
        const uint2 zz2 = *(const uint2*)(sZzc + v * 8u);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const uint32_t pos = ((u < 4 ? zz2.x : zz2.y) >> (8 * (u & 3))) & 0xFFu;
            sA[j * FQ_PITCH + pos] = qv[u];
        }
    }
    __syncthreads();
    // contiguous store of nb * 64 coefficients (scan order per block)
    int32_t* dst = qbuf + ((size_t)P.blk0 + (size_t)T.brow * P.bcols + T.bx0) * 64u;
    const uint32_t total = nb * 64u;
    for (uint32_t i = t * 4u; i < total; i += FQ_THREADS * 4u) {
        const int32_t* s4 = sA + (i >> 6) * FQ_PITCH + (i & 63u);
#if FQ_NT_STORE
        typedef int32_t i32x4 __attribute__((ext_vector_type(4)));
        const i32x4 v4 = { s4[0], s4[1], s4[2], s4[3] };
        __builtin_nontemporal_store(v4, (i32x4*)(dst + i));
#else
        *(int4*)(dst + i) = make_int4(s4[0], s4[1], s4[2], s4[3]);
#endif
    }
}

// grid: x = group of TPW tiles (FQ_BLOCKS blocks of a block row each) within
// the plane, y = plane of the batch; quantiser contexts are indexed by plane,
// so the context and the plane descriptor load in parallel and no tile ->
// plane search is needed. The zigzag table is staged in LDS with the
// quantiser context (a per-lane LDS read is cheaper than a constant-memory
// gather). The samples of all TPW tiles are loaded before the first tile is
// transformed, so a workgroup keeps TPW x 4 KiB of loads in flight.
template <uint32_t TPW>
__global__ __launch_bounds__(FQ_THREADS) void k_fdct_quant(const tcg_qctx* __restrict__ qctxs,
                                                     const tcg_eplane* __restrict__ planes,
                                                     int32_t* __restrict__ qbuf)
{
    __shared__ int32_t sA[FQ_BLOCKS * FQ_PITCH];
    __shared__ uint32_t sQm[64], sQl[64];
    __shared__ uint32_t sBias[64];
    __shared__ __align__(8) uint8_t sZzc[64];  // scan position of natural u * 8 + v at v * 8 + u
    const uint32_t pi = blockIdx.y, t = threadIdx.x;
    const tcg_qctx* qc = &qctxs[pi];
    if (t < 64u) {
        sQm[t] = qc->qmul[t];
        sQl[t] = qc->qsh[t];
        sBias[t] = qc->half[t] - qc->dz[t];
        sZzc[(t & 7u) * 8u + (t >> 3)] = kZigzagInv[t];
    }
    const uint32_t bd = qc->bit_depth;
    const int32_t mid = qc->mid, fclamp = qc->fclamp;
    const tcg_eplane P = planes[pi];
    if (P.kind != 0u) { return; }
    FqTile T[TPW];
    int32_t x[TPW][8];
#pragma unroll
    for (uint32_t k = 0; k < TPW; ++k) {
        T[k] = fq_tile(P, blockIdx.x * TPW + k);
        fq_load(P, T[k], mid, t, x[k]);
    }
#pragma unroll
    for (uint32_t k = 0; k < TPW; ++k) {
        if (!T[k].live) { break; }
        if (k > 0u) { __syncthreads(); }  // the previous tile's store has read sA
        if (bd > 12u) {
            fdct_quant_tile<true>(P, T[k], fclamp, x[k], qbuf, sA, sQm, sQl, sBias, sZzc);
        } else {
            fdct_quant_tile<false>(P, T[k], fclamp, x[k], qbuf, sA, sQm, sQl, sBias, sZzc);
        }
    }
}

// ---------------------------------------------------------------------------
// K2: per-slice rANS2 encoder, one 64-lane wavefront per slice.
// Lane i holds scan position i of the current block. Forward pass: DC
// prediction, tokens and the five joint histograms (LDS atomics). Model
// choice: lanes 0..16 build one candidate row each; five lanes sum the exact
// f64 costs in the reference's order. Backward pass: symbols are pushed in
// reverse decode order with the wave-uniform state; lanes precompute each
// pair's table entries so the serial chain only does readlane + state math.
// ---------------------------------------------------------------------------

// LDS histogram layout (u32)
#define H_DC 0          // 29
#define H_RUN 29        // 64
#define H_LVL 93        // 28
#define H_LPOS 121      // 4 x 28
#define H_LPREV 233     // 5 x 28
#define H_DPREV 373     // 5 x 29
#define H_TOTAL 518

// candidate rows: 0 lvl-none, 1..4 lvl-pos, 5..9 lvl-prev, 10 dc-none,
// 11..15 dc-prev, 16 run
#define NROWS 17

// One symbol's encoder constants, with tc_rans_enc_consts unpacked:
// x' = x + (umulhi(2x, A) >> l) (4096 - f) + cum after renormalising below
// X = f 2^19; X8 = 256 X (saturated: x < 2^31 never reaches it when X >= 2^24)
struct SymK { uint32_t X, X8, A, l, nf, cum; };

// the symbol's constants from a packed (A, B, X) table entry
__device__ static inline SymK sym_k(uint32_t A, uint32_t B, uint32_t X)
{
    SymK k;
    k.X = X; k.X8 = X >= (1u << 24) ? 0xFFFFFFFFu : X << 8;
    k.A = A; k.l = B >> 24; k.nf = (B >> 12) & 0xFFFu; k.cum = B & 0xFFFu;
    return k;
}

// The encoder state stays in [2^23, 2^31). Before a symbol the bytes below
// the state leave until it is under X: 0, 1 or 2 bytes (x >> 16 < 2^15 <=
// X), i.e. a shift of 0, 8 or 16 bits.
__device__ static inline uint32_t enc_sym_shift(uint32_t x, const SymK& k)
{
    return x >= k.X8 ? 16u : (x >= k.X ? 8u : 0u);
}

__device__ static inline uint32_t enc_sym(uint32_t x, const SymK& k)
{
    const uint32_t xr = x >> enc_sym_shift(x, k);
    const uint32_t q = __umulhi(xr << 1, k.A) >> k.l;
    return __umul24(q, k.nf) + xr + k.cum;
}

// Before w raw bits the bytes below the state leave until it is under
// 2^(31 - w): with b = bit length of x (24..31) that is 8 ceil((b + w - 31) / 8)
// bits = (w + 8 - clz(x)) & ~7 (b + w - 24 >= 0 always)
__device__ static inline uint32_t enc_chunk_shift(uint32_t x, uint32_t wp8)
{
    return (wp8 - (uint32_t)__builtin_clz(x)) & 0x38u;
}

// w raw bits c: x' = (x >> shift) 2^w + c; wp8 = w + 8
__device__ static inline uint32_t enc_chunk(uint32_t x, uint32_t wp8, uint32_t w, uint32_t c)
{
    return ((x >> enc_chunk_shift(x, wp8)) << w) | c;
}

// Raw-bit chunks of one position: the suffix of a magnitude category is
// cat - 1 raw bits of the mapped magnitude m; more than 23 bits go as two
// chunks (the top bits, then the low 16), fewer as one (the first chunk is
// then empty: 0 bits, which moves nothing).
struct Chunks { uint32_t wp8a, wa, ca, wp8b, wb, cb; };

__device__ static inline Chunks make_chunks(uint32_t w, uint32_t m)
{
    Chunks c;
    if (w > 23u) {
        c.wa = w - 16u; c.ca = (m >> 16) & ((1u << (w - 16u)) - 1u);
        c.wb = 16u; c.cb = m & 0xFFFFu;
    } else {
        c.wa = 0u; c.ca = 0u;
        c.wb = w; c.cb = m & ((1u << w) - 1u);
    }
    c.wp8a = c.wa + 8u; c.wp8b = c.wb + 8u;
    return c;
}

// The serial chain over one block's active positions (bits of m, highest
// first). Every lane evaluates position p's operations with its own
// constants (the run symbol's renormalisation, which only depends on the
// broadcast state and lane p's bound, runs on the scalar unit); lane p's
// results are the real ones: it logs the states before and after its level
// symbol (and, WIDE, before its second chunk) for the byte pass, and its final
// state is broadcast to the next position. WIDE: a first chunk of
// some lane is not empty.
template <bool WIDE>
__device__ static inline uint32_t enc_positions(uint32_t x, uint64_t m, const SymK& rk, const Chunks& ch,
                                                const SymK& bk, uint32_t* sLog, uint32_t lane)
{
    uint32_t* const mylog = sLog + 4u * lane;   // lane p writes entry p
    do {
        const uint32_t p = 63u - (uint32_t)__clzll(m);
        m &= ~(1ull << p);
        // the run symbol: its renormalisation on the scalar unit with lane
        // p's bound (the state is wave-uniform), the update in every lane
        const uint32_t X = __builtin_amdgcn_readlane(rk.X, p);
        const uint32_t x1 = x >= X ? x >> 8 : x;
        const uint32_t xr = x1 >= X ? x1 >> 8 : x1;
        const uint32_t xa = __umul24(__umulhi(xr << 1, rk.A) >> rk.l, rk.nf) + xr + rk.cum;
        const uint32_t xa2 = WIDE ? enc_chunk(xa, ch.wp8a, ch.wa, ch.ca) : xa;
        const uint32_t xb = enc_chunk(xa2, ch.wp8b, ch.wb, ch.cb);
        const uint32_t xo = enc_sym(xb, bk);
        if (lane == p) {
            if (WIDE) { *(uint4*)mylog = make_uint4(xb, xo, xa2, 0u); }
            else { *(uint2*)mylog = make_uint2(xb, xo); }
        }
        x = __builtin_amdgcn_readlane(xo, p);
    } while (m != 0ull);
    return x;
}

__device__ static inline uint32_t wave_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

#define ENC_GROUP 8
// words of the model tables: row bytes, frequencies and cumulative frequencies
#define ENC_MODEL_WORDS ((NROWS * 64 + NROWS * 64 * 2 + NROWS * 65 * 2 + 3) / 4)
// five f64 costs (one spare) and the slice prefix bytes
#define ENC_FIXED_WORDS ((12 + (TCG_PREFIX_MAX + 3) / 4 + 3) & ~3)

// encoder constants of the backward pass (three tables of 11 x 64 words) and
// its per-position state log (64 x 4 words, 16-byte aligned: the states
// before and after the level symbol and before a second chunk)
#define ENC_TAB_WORDS (3u * 11u * 64u)
#define ENC_LOG_WORDS (4u * 64u)

// LDS words of one encoding wave whose plane is `cols` blocks wide: the phase
// space (histograms + DC row | model tables | encoder constants + state log),
// then the fixed part; a multiple of four words, so every wave's space is
// 16-byte aligned
__host__ __device__ static inline uint32_t enc_phase_words(uint32_t cols)
{
    const uint32_t a = H_TOTAL + (cols > ENC_MODEL_WORDS ? cols : ENC_MODEL_WORDS);
    const uint32_t b = ENC_TAB_WORDS + ENC_LOG_WORDS;
    return ((a > b ? a : b) + 3u) & ~3u;  // the costs that follow are f64
}

__device__ static void slice_encode_one(const TcgConsts* __restrict__ K, const tcg_eplane* __restrict__ planes,
                                        const tcg_eslice* __restrict__ slices, const int32_t* __restrict__ qbuf,
                                        uint32_t* __restrict__ dcmbuf, uint8_t* __restrict__ scratch,
                                        tcg_eresult* __restrict__ results, uint32_t si, uint32_t lane,
                                        uint32_t* __restrict__ sDyn)
{
    // this wave's LDS (words), reused across the slice's phases:
    //  forward pass      [0, H_TOTAL) histograms, then the DC values of the
    //                    block row above (one per block column of the batch's
    //                    widest plane)
    //  model choice      histograms, then the rows' bytes, frequencies and
    //                    cumulative frequencies over the dead DC row
    //  backward pass     [0, 3 x 11 x 64) per-symbol encoder constants, stored
    //                    from registers after the model tables' last read,
    //                    then the per-position state log (64 x 4 words)
    //  always            the five model costs and the slice prefix bytes, in
    //                    the ENC_FIXED_WORDS after the phase space
    uint32_t* const sH = sDyn;
    int32_t* const sDcRow = (int32_t*)(sDyn + H_TOTAL);
    uint8_t (*const sRowB)[64] = (uint8_t (*)[64])(sDyn + H_TOTAL);
    uint16_t (*const sFreq)[64] = (uint16_t (*)[64])((uint8_t*)(sDyn + H_TOTAL) + NROWS * 64);
    uint16_t (*const sCum)[65] = (uint16_t (*)[65])((uint8_t*)(sDyn + H_TOTAL) + NROWS * 64 * 3);
    uint32_t* const sTabA = sDyn;
    uint32_t* const sTabB = sDyn + 11 * 64;
    uint32_t* const sTabX = sDyn + 2 * 11 * 64;
    const tcg_eslice S = slices[si];
    const uint32_t cols = planes[S.plane].bcols;
    double* const sCost = (double*)(sDyn + enc_phase_words(cols));
    uint8_t* const sPrefix = (uint8_t*)(sCost + 6);
    if (S.mode != TCG_MODE_RANS2) { return; }
    const tcg_eplane P = planes[S.plane];
    if (P.kind != 0u) { return; }

    for (uint32_t i = lane; i < H_TOTAL; i += 64u) { sH[i] = 0u; }
    wave_sync();

    // ---------------- forward pass ----------------
    // One wavefront: LDS accesses of this wave complete in program order, so
    // the row of DC values and the histograms need no barrier per block.
    uint32_t viol = 0u;
    uint32_t prev_dcc = 0u;
    int32_t left = 0;
    const uint64_t below = lane == 0u ? 0ull : ((1ull << lane) - 1ull);
    const uint32_t nblk = S.nblk;
    const int32_t* qs = qbuf + (size_t)S.blk0 * 64u + lane;
    uint32_t bx = 0u, by = 0u;
    auto fwd_block = [&](int32_t q, uint32_t b) {
        const uint64_t nz = __ballot(q != 0) & ~1ull;
        const int32_t q0 = __builtin_amdgcn_readlane(q, 0);
        const int32_t top = by ? (int32_t)wave_u32((uint32_t)sDcRow[bx]) : 0;
        const int32_t pred = tc_dc_predict(bx != 0u, left, by != 0u, top);
        uint32_t dcm = tc_map_signed(q0 - pred);
        if (dcm > 0x0FFFFFFFu) { dcm = 0x0FFFFFFFu; viol = 1u; }
        const uint32_t dcc = tc_bitlen(dcm);
        if (lane == 0u) {
            dcmbuf[S.blk0 + b] = dcm;
            atomicAdd(&sH[H_DC + dcc], 1u);
            atomicAdd(&sH[H_DPREV + (b == 0u ? 0u : tc_ctx_dc(prev_dcc, 1u)) * 29u + dcc], 1u);
            atomicAdd(&sH[H_RUN + 63u], 1u);
            sDcRow[bx] = q0;
        }
        prev_dcc = dcc;
        left = q0;
        const uint32_t lm_raw = tc_map_signed(q);
        const uint32_t lm = min(lm_raw, 0x07FFFFFFu);
        const uint32_t cat = tc_bitlen(lm);
        const uint64_t bl = nz & below;
        const uint32_t pp = bl ? 63u - (uint32_t)__clzll(bl) : 0u;
        const uint32_t pcat = (uint32_t)__shfl((int)cat, (int)pp);
        if (lane != 0u && q != 0) {
            if (lm_raw > 0x07FFFFFFu) { viol = 1u; }
            const uint32_t run = lane - pp - 1u;
            atomicAdd(&sH[H_RUN + run], 1u);
            atomicAdd(&sH[H_LVL + cat], 1u);
            atomicAdd(&sH[H_LPOS + tc_ctx_pos(lane) * 28u + cat], 1u);
            atomicAdd(&sH[H_LPREV + tc_ctx_prev(pcat, pp != 0u) * 28u + cat], 1u);
        }
        if (++bx == cols) { bx = 0u; ++by; }
    };
    // groups of ENC_GROUP blocks, the next group's loads in flight while the
    // current one is processed
    int32_t qa[ENC_GROUP], qb[ENC_GROUP];
#pragma unroll
    for (uint32_t j = 0; j < ENC_GROUP; ++j) { qa[j] = j < nblk ? qs[(size_t)j * 64u] : 0; }
    for (uint32_t g = 0; g < nblk; g += ENC_GROUP) {
#pragma unroll
        for (uint32_t j = 0; j < ENC_GROUP; ++j) {
            const uint32_t b = g + ENC_GROUP + j;
            qb[j] = b < nblk ? qs[(size_t)b * 64u] : 0;
        }
#pragma unroll
        for (uint32_t j = 0; j < ENC_GROUP; ++j) {
            if (g + j < nblk) { fwd_block(qa[j], g + j); }
        }
#pragma unroll
        for (uint32_t j = 0; j < ENC_GROUP; ++j) { qa[j] = qb[j]; }
    }
    wave_sync();
    const uint32_t any_viol = __ballot(viol != 0u) != 0ull;

    // ---------------- model choice ----------------
    // candidate rows one at a time, one symbol per lane (row bytes from the
    // counts, then the model; rows from nonzero bytes always fit)
    for (uint32_t r = 0; r < NROWS; ++r) {
        const uint32_t* cnt;
        uint32_t nsym;
        if (r == 0u) { cnt = sH + H_LVL; nsym = 28u; }
        else if (r <= 4u) { cnt = sH + H_LPOS + (r - 1u) * 28u; nsym = 28u; }
        else if (r <= 9u) { cnt = sH + H_LPREV + (r - 5u) * 28u; nsym = 28u; }
        else if (r == 10u) { cnt = sH + H_DC; nsym = 29u; }
        else if (r <= 15u) { cnt = sH + H_DPREV + (r - 11u) * 29u; nsym = 29u; }
        else { cnt = sH + H_RUN; nsym = 64u; }
        const uint32_t b = wave_row_byte(lane < nsym ? cnt[lane] : 0u, lane);
        uint32_t f = 0u, c = 0u;
        (void)wave_model_from_bytes(b, lane, f, c);
        if (lane < nsym) {
            sRowB[r][lane] = (uint8_t)b;
            sFreq[r][lane] = (uint16_t)f;
            sCum[r][lane] = (uint16_t)c;
        }
        if (lane == 0u) { sCum[r][nsym] = 4096u; }
    }
    wave_sync();
    // exact cost sums, one candidate per lane, in the reference's order:
    // lane 0 lvl-none, 1 lvl-pos, 2 lvl-prev, 3 dc-none, 4 dc-prev
    if (lane < 5u) {
        uint32_t r0, nr, nsym;
        const uint32_t* base;
        if (lane == 0u) { r0 = 0; nr = 1; nsym = 28; base = sH + H_LVL; }
        else if (lane == 1u) { r0 = 1; nr = 4; nsym = 28; base = sH + H_LPOS; }
        else if (lane == 2u) { r0 = 5; nr = 5; nsym = 28; base = sH + H_LPREV; }
        else if (lane == 3u) { r0 = 10; nr = 1; nsym = 29; base = sH + H_DC; }
        else { r0 = 11; nr = 5; nsym = 29; base = sH + H_DPREV; }
        double bits = 0.0;
        for (uint32_t r = 0; r < nr; ++r) {
            for (uint32_t s = 0; s < nsym; ++s) {
                const uint32_t cn = base[r * nsym + s];
                if (cn != 0u) { bits = cost_add(K, bits, cn, sFreq[r0 + r][s]); }
            }
        }
        sCost[lane] = cost_add_const(bits, 8.0 * (double)(nr * nsym));
    }
    wave_sync();
    uint32_t lvl_model = 0u, dc_model = 0u;
    {
        double best = sCost[0];
        if (sCost[1] < best) { best = sCost[1]; lvl_model = 1u; }
        if (sCost[2] < best) { lvl_model = 2u; }
        if (sCost[4] < sCost[3]) { dc_model = 1u; }
    }
    lvl_model = wave_u32(lvl_model);
    dc_model = wave_u32(dc_model);
    const uint32_t lvl_n = lvl_model == 1u ? 4u : (lvl_model == 2u ? 5u : 1u);
    const uint32_t dc_n = dc_model ? 5u : 1u;
    const uint32_t lvl_r0 = lvl_model == 1u ? 1u : (lvl_model == 2u ? 5u : 0u);
    const uint32_t dc_r0 = dc_model ? 11u : 10u;
    // table slots: [0, dc_n) dc, dc_n run, then lvl rows
    const uint32_t prefix_len = 1u + dc_n * 29u + 64u + lvl_n * 28u;
    {
        // prefix bytes
        if (lane == 0u) { sPrefix[0] = (uint8_t)(lvl_model | (dc_model << 2)); }
        for (uint32_t i = lane; i < dc_n * 29u; i += 64u) {
            sPrefix[1u + i] = sRowB[dc_r0 + i / 29u][i % 29u];
        }
        sPrefix[1u + dc_n * 29u + lane] = sRowB[16][lane];
        for (uint32_t i = lane; i < lvl_n * 28u; i += 64u) {
            sPrefix[1u + dc_n * 29u + 64u + i] = sRowB[lvl_r0 + i / 28u][i % 28u];
        }
    }
    {
        // per-symbol encoder constants: computed into registers while the
        // model tables are live, stored over them after the barrier
        uint32_t tA[11], tB[11], tX[11];
        const uint32_t nslot = dc_n + 1u + lvl_n;
#pragma unroll
        for (uint32_t slot = 0; slot < 11u; ++slot) {
            uint32_t row, nsym;
            if (slot < dc_n) { row = dc_r0 + slot; nsym = 29u; }
            else if (slot == dc_n) { row = 16u; nsym = 64u; }
            else { row = lvl_r0 + (slot - dc_n - 1u); nsym = 28u; }
            tA[slot] = tB[slot] = tX[slot] = 0u;
            if (slot < nslot && lane < nsym) {
                tc_rans_enc_consts(sFreq[row][lane], sCum[row][lane], &tA[slot], &tB[slot], &tX[slot]);
            }
        }
        wave_sync();
#pragma unroll
        for (uint32_t slot = 0; slot < 11u; ++slot) {
            if (slot < nslot) {
                sTabA[slot * 64u + lane] = tA[slot];
                sTabB[slot * 64u + lane] = tB[slot];
                sTabX[slot * 64u + lane] = tX[slot];
            }
        }
    }
    wave_sync();

    // ---------------- backward pass ----------------
    // Symbols are pushed in reverse decode order. Per block, every position
    // holding a nonzero level, and position 0 (the DC), pushes three things:
    // the run symbol decoded after it (the step to the next such position
    // above, or end of block), its suffix chunk(s), and its level symbol (the
    // DC category symbol at position 0). Lane p holds position p's constants;
    // the serial chain runs position p's operations in every lane at once from
    // the broadcast state and keeps lane p's result (one readlane per
    // position), so the chain is vector arithmetic with no branches. Bytes are
    // not emitted on the chain: lane p logs the states before its operations,
    // and after the block every lane recomputes its renormalisation shifts,
    // finds where its bytes go from the byte counts of the positions above it
    // and stores them.
    const uint32_t run_slot = dc_n;
    const uint32_t lvl_slot0 = dc_n + 1u;
    uint32_t* const sLog = sDyn + ENC_TAB_WORDS;
    uint8_t* const end = scratch + S.scratch + S.scratch_cap;
    const uint32_t limit = S.scratch_cap - TCG_PREFIX_MAX - 8u;
    uint32_t n = 0u, over = 0u;
    uint32_t x = 1u << 23;
    const uint64_t above = lane == 63u ? 0ull : (~0ull << (lane + 1u));
    // one block ahead: its coefficients and the DC value of the block after it
    int32_t q_nx = nblk ? qs[(size_t)(nblk - 1u) * 64u] : 0;
    uint32_t d_cur = nblk ? dcmbuf[S.blk0 + nblk - 1u] : 0u;
    uint32_t d_nx = nblk >= 2u ? dcmbuf[S.blk0 + nblk - 2u] : 0u;
    for (uint32_t bb = nblk; bb-- > 0u;) {
        const int32_t q = q_nx;
        const uint32_t dcm = wave_u32(d_cur);
        const uint32_t dcm_prev = wave_u32(d_nx);
        if (bb >= 1u) { q_nx = qs[(size_t)(bb - 1u) * 64u]; }
        d_cur = d_nx;
        d_nx = bb >= 2u ? dcmbuf[S.blk0 + bb - 2u] : 0u;
        const uint64_t nz = __ballot(q != 0) & ~1ull;
        const uint64_t act = nz | 1ull;
        const uint32_t lm = min(tc_map_signed(q), 0x07FFFFFFu);
        const uint32_t cat = tc_bitlen(lm);
        const uint64_t bl = nz & below;
        const uint32_t pp = bl ? 63u - (uint32_t)__clzll(bl) : 0u;
        const uint32_t pcat = (uint32_t)__shfl((int)cat, (int)pp);
        uint32_t lc = 0u;
        if (lvl_model == 1u) { lc = tc_ctx_pos(lane); }
        else if (lvl_model == 2u) { lc = tc_ctx_prev(pcat, pp != 0u); }
        const bool is_act = ((act >> lane) & 1ull) != 0ull;
        // the next active position above: the run decoded after this one
        const uint64_t ab = act & above;
        const uint32_t nx = ab ? (uint32_t)__builtin_ctzll(ab) : 64u;
        const uint32_t rsym = nx == 64u ? 63u : nx - lane - 1u;
        // position 0: the DC category in its context row, its suffix
        const uint32_t dsym = tc_bitlen(dcm);
        const uint32_t dctx = dc_model ? tc_ctx_dc(tc_bitlen(dcm_prev), bb != 0u) : 0u;
        const uint32_t bi = lane == 0u ? dctx * 64u + dsym : (is_act ? (lvl_slot0 + lc) * 64u + cat : 0u);
        const uint32_t ri = is_act ? run_slot * 64u + rsym : 0u;
        const SymK rk = sym_k(sTabA[ri], sTabB[ri], sTabX[ri]);
        const SymK bk = sym_k(sTabA[bi], sTabB[bi], sTabX[bi]);
        const uint32_t ccat = lane == 0u ? dsym : cat;
        const Chunks ch = make_chunks(ccat > 1u ? ccat - 1u : 0u, lane == 0u ? dcm : lm);
        const uint32_t x_in = x;
        const bool wide = __ballot(is_act && ch.wa != 0u) != 0ull;
        if (!wide) {
            x = enc_positions<false>(x, act, rk, ch, bk, sLog, lane);
        } else {
            x = enc_positions<true>(x, act, rk, ch, bk, sLog, lane);
        }
        // this block's bytes, in push order: positions from the top, per
        // position the run symbol's, the chunks' and the level symbol's
        uint32_t nb = 0u, kR = 0u, kRA = 0u, kRA2 = 0u;
        uint32_t xr = 0u, xa = 0u, xa2 = 0u, xb = 0u;
        if (is_act) {
            const uint4 e = *(const uint4*)(sLog + 4u * lane);
            xb = e.x;
            xr = nx == 64u ? x_in : sLog[4u * nx + 1u];
            xa = enc_sym(xr, rk);
            xa2 = wide ? e.z : xa;
            kR = enc_sym_shift(xr, rk) >> 3;
            kRA = kR + (enc_chunk_shift(xa, ch.wp8a) >> 3);
            kRA2 = kRA + (enc_chunk_shift(xa2, ch.wp8b) >> 3);
            nb = kRA2 + (enc_sym_shift(xb, bk) >> 3);
        }
        // bytes before this lane's: those of the positions above (nb <= 10)
        const uint64_t b0 = __ballot(nb & 1u), b1 = __ballot(nb & 2u), b2 = __ballot(nb & 4u),
                       b3 = __ballot(nb & 8u);
        const uint32_t tot = (uint32_t)(__popcll(b0) + 2 * __popcll(b1) + 4 * __popcll(b2) + 8 * __popcll(b3));
        if (n + tot > limit) { over = 1u; break; }
        if (tot != 0u) {
            const uint32_t off = (uint32_t)(__popcll(b0 & above) + 2 * __popcll(b1 & above) +
                                            4 * __popcll(b2 & above) + 8 * __popcll(b3 & above));
            // byte j of this lane is the (n + off + j + 1)-th byte pushed
            uint8_t* const base = end - (size_t)(n + off) - 1;
            for (uint32_t j = 0u; __ballot(j < nb) != 0ull; ++j) {
                const uint32_t src = j < kR ? xr : (j < kRA ? xa : (j < kRA2 ? xa2 : xb));
                const uint32_t sft = j - (j < kR ? 0u : (j < kRA ? kR : (j < kRA2 ? kRA : kRA2)));
                if (j < nb) { *(base - j) = (uint8_t)(src >> (8u * sft)); }
            }
            n += tot;
        }
    }
    // ---------------- flush ----------------
    // the pushed bytes occupy [end - n, end); the final state (big-endian)
    // and the prefix go in front of them
    const uint32_t ok_len = (over == 0u && n <= limit) ? 1u : 0u;
    if (ok_len) {
        uint8_t* st = end - n - 4;
        if (lane < 4u) { st[lane] = (uint8_t)(x >> (24u - 8u * lane)); }
        uint8_t* pre = st - prefix_len;
        for (uint32_t i = lane; i < prefix_len; i += 64u) { pre[i] = sPrefix[i]; }
    }
    if (lane == 0u) {
        tcg_eresult R;
        R.start = ok_len ? S.scratch_cap - n - 4u - prefix_len : 0u;
        R.size = ok_len ? n + 4u + prefix_len : 0u;
        R.status = any_viol ? -1 : (ok_len ? 0 : -15);
        R.k1 = 0; R.k2 = 0; R.k3 = 0; R.crc = 0; R.offset = 0; R.aux = 0; R.pad_ = 0;
        results[si] = R;
    }
}

// Persistent launch: the batch's rANS2 colour slices sit in a work list,
// longest first; wave k of the grid encodes entries k, k + W, ... (W waves in
// the grid), so a grid smaller than the list makes every wave work through
// several slices. A workgroup holds `wpg` waves, each with its own LDS of
// `wave_words` words.
__global__ __launch_bounds__(256) void k_slice_encode(const TcgConsts* __restrict__ K,
                                                       const tcg_eplane* __restrict__ planes,
                                                       const tcg_eslice* __restrict__ slices,
                                                       const int32_t* __restrict__ qbuf,
                                                       uint32_t* __restrict__ dcmbuf,
                                                       uint8_t* __restrict__ scratch,
                                                       tcg_eresult* __restrict__ results,
                                                       const tcg_work* __restrict__ work, uint32_t wave_words)
{
    extern __shared__ uint32_t sWaves[];
    const uint32_t lane = threadIdx.x & 63u;
    const WaveSlot ws = wave_slot();
    uint32_t* const mine = sWaves + ws.wave * wave_words;
    const uint32_t* const order = tcg_work_order(work);
    const uint32_t count = work->count;
    for (uint32_t i = ws.first; i < count; i += ws.stride) {
        slice_encode_one(K, planes, slices, qbuf, dcmbuf, scratch, results, order[i], lane, mine);
    }
}

// ---------------------------------------------------------------------------
// K3: per-frame offset scan + frame header (53 B, CRC32 over bytes 0..48).
// One wave per frame: lane l sums the slices of its share, a wave prefix sum
// gives every share its offset. A one-wave workgroup: when another launch
// sequence's entropy kernel holds most of the device, any single freed wave
// slot can take it, where a four-wave workgroup would wait for four slots and
// their LDS on one CU (measured: up to 25 ms of waiting for a 0.04 ms
// kernel).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_frame_scan(const TcgConsts* __restrict__ K,
                                                    const tcg_eframe* __restrict__ frames,
                                                    tcg_eresult* __restrict__ results,
                                                    const tcg_eslice* __restrict__ slices,
                                                    const tcg_eplane* __restrict__ planes,
                                                    tcg_eframe_out* __restrict__ fout)
{
    const tcg_eframe F = frames[blockIdx.x];
    const uint32_t lane = threadIdx.x;
    const uint32_t n = F.s1 - F.s0;
    const uint32_t per = (n + 63u) / 64u;
    const uint32_t a = F.s0 + min(n, lane * per), b = F.s0 + min(n, (lane + 1u) * per);
    uint32_t sum = 0u, col = 0u, al = 0u, coln = 0u, aln = 0u;
    int32_t stat = 0;
    for (uint32_t s = a; s < b; ++s) {
        sum += 17u + results[s].size;
        if (results[s].status != 0 && stat == 0) { stat = results[s].status; }
        if (planes[slices[s].plane].kind == 1u) { al += results[s].size; aln++; }
        else { col += results[s].size; coln++; }
    }
    // inclusive prefix of the shares' byte counts, then this share's offset
    uint32_t inc = sum;
#pragma unroll
    for (uint32_t d = 1u; d < 64u; d <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)inc, d);
        if (lane >= d) { inc += v; }
    }
    const uint32_t total = 53u + (uint32_t)__shfl((int)inc, 63);
    const uint32_t c = wave_sum(col), l = wave_sum(al), cn = wave_sum(coln), ln = wave_sum(aln);
    // the first failing slice's status, in slice order
    const uint64_t bad = __ballot(stat != 0);
    int32_t st = bad ? __shfl(stat, (int)__builtin_ctzll(bad)) : 0;
    if (st == 0 && total > F.cap) { st = -15; }
    if (st == 0 && total > 268435456u) { st = -8; }
    if (lane == 0u) {
        tcg_eframe_out o;
        o.size = total; o.status = st; o.color_payload = c; o.alpha_payload = l;
        o.color_slices = cn; o.alpha_slices = ln;
        fout[blockIdx.x] = o;
        if (st == 0) {
            uint8_t h[53];
            for (int i = 0; i < 53; ++i) { h[i] = F.hdr[i]; }
            h[41] = (uint8_t)(total >> 24); h[42] = (uint8_t)(total >> 16);
            h[43] = (uint8_t)(total >> 8); h[44] = (uint8_t)total;
            const uint32_t crc = crc32_bytes_tab(K->crc, h, 49u);
            h[49] = (uint8_t)(crc >> 24); h[50] = (uint8_t)(crc >> 16);
            h[51] = (uint8_t)(crc >> 8); h[52] = (uint8_t)crc;
            uint8_t* pk = (uint8_t*)F.packet;
            for (int i = 0; i < 53; ++i) { pk[i] = h[i]; }
        }
    }
    uint32_t off = 53u + inc - sum;
    for (uint32_t s = a; s < b; ++s) {
        results[s].offset = off + 17u;
        off += 17u + results[s].size;
    }
}

// ---------------------------------------------------------------------------
// K4: per-slice payload copy into the packet + CRC32 + 17-byte slice header.
// The payload moves in 4 KiB tiles aligned to the destination: the source
// bytes of a tile are staged in LDS with coalesced 16-byte loads, then thread
// t owns the t-th aligned 16-byte word of the tile in the packet: it writes
// the word (one 16-byte store when the word is whole) and computes the raw
// CRC of its bytes, shifted to the tile end by one GF(2) multiplication
// (x^(128 j) table; the last tile uses the general shift). The XOR of the
// shifted word CRCs is the tile's raw CRC; tiles chain by a shift over their
// length. The standard init / final XOR is applied once at the end.
// ---------------------------------------------------------------------------
#define EMIT_TILE 4096u

__global__ __launch_bounds__(256) void k_slice_emit(const TcgConsts* __restrict__ K,
                                                     const tcg_eframe* __restrict__ frames,
                                                     const tcg_eslice* __restrict__ slices,
                                                     const uint8_t* __restrict__ scratch,
                                                     tcg_eresult* __restrict__ results,
                                                     const tcg_eframe_out* __restrict__ fout)
{
    __shared__ uint32_t sTab[256];
    __shared__ uint32_t sX16[256];
    __shared__ uint32_t sX2n[32];
    __shared__ uint4 sBuf[EMIT_TILE / 16u + 2u];
    __shared__ uint32_t sRed[4];
    const uint32_t s = blockIdx.x, t = threadIdx.x;
    const tcg_eslice S = slices[s];
    const tcg_eresult R = results[s];
    const tcg_eframe F = frames[S.frame];
    sTab[t] = K->crc[t];
    sX16[t] = K->x16[t];
    if (t < 32u) { sX2n[t] = K->x2n[t]; }
    __syncthreads();
    if (fout[S.frame].status != 0) { return; }
    const uint8_t* src = scratch + S.scratch + R.start;
    uint8_t* dst = (uint8_t*)F.packet + R.offset;
    const uint32_t n = R.size;
    const uintptr_t d0 = (uintptr_t)dst, dend = d0 + n;
    uint32_t craw = 0u;  // raw CRC of the bytes emitted so far (same in every thread)
    for (uintptr_t T = d0 & ~(uintptr_t)(EMIT_TILE - 1u); T < dend; T += EMIT_TILE) {
        const uintptr_t ts = T > d0 ? T : d0;
        const uintptr_t te = T + EMIT_TILE < dend ? T + EMIT_TILE : dend;
        const uint32_t L = (uint32_t)(te - ts);
        // stage source bytes [ts - d0, te - d0)
        const uintptr_t s0 = (uintptr_t)(src + (ts - d0));
        const uintptr_t sa = s0 & ~(uintptr_t)15u;
        const uint32_t soff = (uint32_t)(s0 - sa);
        const uint32_t nvec = (soff + L + 15u) >> 4;
        for (uint32_t i = t; i < nvec; i += 256u) { sBuf[i] = ((const uint4*)sa)[i]; }
        __syncthreads();
        const uint8_t* sb = (const uint8_t*)sBuf + soff;
        const uintptr_t W = (ts & ~(uintptr_t)15u) + 16u * t;
        uint32_t part = 0u;
        if (W < te) {
            const uintptr_t a = W > ts ? W : ts;
            const uintptr_t b = W + 16u < te ? W + 16u : te;
            const uint32_t o = (uint32_t)(a - ts), len = (uint32_t)(b - a);
            uint32_t c = 0u;
            if (len == 16u) {
                uint32_t w4[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    uint32_t v = 0u;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t byte = sb[o + 4u * k + j];
                        c = sTab[(c ^ byte) & 0xFFu] ^ (c >> 8);
                        v |= byte << (8 * j);
                    }
                    w4[k] = v;
                }
                *(uint4*)W = make_uint4(w4[0], w4[1], w4[2], w4[3]);
            } else {
                for (uint32_t j = 0; j < len; ++j) {
                    const uint8_t byte = sb[o + j];
                    c = sTab[(c ^ byte) & 0xFFu] ^ (c >> 8);
                    ((uint8_t*)a)[j] = byte;
                }
            }
            const uint32_t dist = (uint32_t)(te - b);
            if ((te & 15u) == 0u) {
                part = tc_gf2_mulmod(sX16[dist >> 4], c);
            } else {
                part = tc_crc_shift(sX2n, c, dist);
            }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) { part ^= (uint32_t)__shfl_xor((int)part, off); }
        if ((t & 63u) == 0u) { sRed[t >> 6] = part; }
        __syncthreads();
        const uint32_t tcrc = sRed[0] ^ sRed[1] ^ sRed[2] ^ sRed[3];
        craw = (L == EMIT_TILE ? tc_gf2_mulmod(sX2n[15], craw) : tc_crc_shift(sX2n, craw, L)) ^ tcrc;
        __syncthreads();
    }
    if (t == 0u) {
        const uint32_t crc = craw ^ tc_crc_shift(sX2n, 0xFFFFFFFFu, n) ^ 0xFFFFFFFFu;
        results[s].crc = crc;
        uint8_t* h = dst - 17;
        h[0] = (uint8_t)(n >> 24); h[1] = (uint8_t)(n >> 16); h[2] = (uint8_t)(n >> 8); h[3] = (uint8_t)n;
        h[4] = (uint8_t)S.hdr_plane;
        h[5] = (uint8_t)(S.y0 >> 8); h[6] = (uint8_t)S.y0;
        h[7] = (uint8_t)(S.h >> 8); h[8] = (uint8_t)S.h;
        h[9] = (uint8_t)S.qpd;
        h[10] = (uint8_t)R.k1; h[11] = (uint8_t)R.k2; h[12] = (uint8_t)R.k3;
        h[13] = (uint8_t)(crc >> 24); h[14] = (uint8_t)(crc >> 16); h[15] = (uint8_t)(crc >> 8);
        h[16] = (uint8_t)crc;
    }
}

// symbol pairs of the batch's rANS2 colour slices (timed sequences only)
__global__ __launch_bounds__(256) void k_count_pairs_enc(const tcg_eplane* __restrict__ planes,
                                                         const tcg_eslice* __restrict__ slices,
                                                         const int32_t* __restrict__ q, unsigned long long* counter)
{
    const tcg_eslice S = slices[blockIdx.x];
    if (S.mode != TCG_MODE_RANS2 || planes[S.plane].kind != 0u) { return; }
    tc_count_pairs(q, S.blk0, S.nblk, counter);
}

// ---------------------------------------------------------------------------
// host launch
// ---------------------------------------------------------------------------
extern "C" int32_t tcg_encode(tcg_ctx* c, const tcg_qctx* qctx_dev, uint32_t nplanes, const tcg_eplane* planes_dev,
                              uint32_t nslices, const tcg_eslice* slices_dev, uint32_t nframes,
                              const tcg_eframe* frames_dev, uint64_t qbuf, uint64_t dcmbuf,
                              uint32_t max_plane_tiles, uint64_t scratch, tcg_eresult* results_dev,
                              tcg_eframe_out* frame_out_dev, const tcg_batch_info* info)
{
    hipStream_t st = c->stream;
    const TcgConsts* K = tcg_consts_dev();
    c->kind = 1;
    const uint32_t wpg = info->ent_wpg ? info->ent_wpg : 1u;
    const uint32_t fq_tpw = info->tile_tpw ? info->tile_tpw : 1u;
    const uint32_t wave_words = enc_phase_words(info->max_cols) + ENC_FIXED_WORDS;
    const uint32_t groups = info->ent_groups ? info->ent_groups : (info->ent_count + wpg - 1u) / wpg;
    auto launch = [&]() {
        tcg_mark(c, 0);
        if (max_plane_tiles > 0u) {
            if (fq_tpw == 4u) {
                hipLaunchKernelGGL(k_fdct_quant<4>, dim3((max_plane_tiles + 3u) / 4u, nplanes), dim3(FQ_THREADS), 0,
                                   st, qctx_dev, planes_dev, (int32_t*)qbuf);
            } else if (fq_tpw == 2u) {
                hipLaunchKernelGGL(k_fdct_quant<2>, dim3((max_plane_tiles + 1u) / 2u, nplanes), dim3(FQ_THREADS), 0,
                                   st, qctx_dev, planes_dev, (int32_t*)qbuf);
            } else {
                hipLaunchKernelGGL(k_fdct_quant<1>, dim3(max_plane_tiles, nplanes), dim3(FQ_THREADS), 0, st,
                                   qctx_dev, planes_dev, (int32_t*)qbuf);
            }
        }
        tcg_mark(c, 1);
        if ((info->kinds & TCG_HAS_RANS2) && groups > 0u) {
            hipLaunchKernelGGL(k_slice_encode, dim3(groups), dim3(64u * wpg), (size_t)wave_words * 4u * wpg, st, K,
                               planes_dev, slices_dev, (const int32_t*)qbuf, (uint32_t*)dcmbuf, (uint8_t*)scratch,
                               results_dev, info->work, wave_words);
        }
        if (info->kinds & TCG_HAS_BITS) {
            tcg_bits_encode_launch(st, K, planes_dev, slices_dev, nslices, (const int32_t*)qbuf,
                                   (uint32_t*)dcmbuf, (uint8_t*)scratch, results_dev, info->max_cols);
        }
        if (info->kinds & TCG_HAS_ALPHA) {
            tcg_alpha_encode_launch(st, planes_dev, slices_dev, nslices, nullptr, (uint8_t*)scratch, results_dev);
        }
        tcg_mark(c, 2);
        hipLaunchKernelGGL(k_frame_scan, dim3(nframes), dim3(64), 0, st, K, frames_dev, results_dev,
                           slices_dev, planes_dev, frame_out_dev);
        tcg_mark(c, 3);
        hipLaunchKernelGGL(k_slice_emit, dim3(nslices), dim3(256), 0, st, K, frames_dev, slices_dev,
                           (const uint8_t*)scratch, results_dev, frame_out_dev);
        tcg_mark(c, 4);
        unsigned long long* pairs = (info->kinds & TCG_HAS_RANS2) ? tcg_pairs_counter(c) : nullptr;
        if (pairs != nullptr) {
            hipLaunchKernelGGL(k_count_pairs_enc, dim3(nslices), dim3(256), 0, st, planes_dev, slices_dev,
                               (const int32_t*)qbuf, pairs);
        }
    };
    const uint64_t key[TCG_GRAPH_KEY] = {
        1u, (uint64_t)(uintptr_t)qctx_dev, nplanes, (uint64_t)(uintptr_t)planes_dev, nslices,
        (uint64_t)(uintptr_t)slices_dev, nframes, (uint64_t)(uintptr_t)frames_dev, qbuf, dcmbuf, max_plane_tiles,
        scratch, (uint64_t)(uintptr_t)results_dev, (uint64_t)(uintptr_t)frame_out_dev, info->max_cols, info->kinds,
        (uint64_t)(uintptr_t)K, (uint64_t)wpg << 32 | groups, (uint64_t)(uintptr_t)info->work, fq_tpw };
    return tcg_run(c, key, launch);
}

uint32_t tcg_entropy_capacity_enc(uint32_t wpg, uint32_t max_cols)
{
    const uint32_t wave_words = enc_phase_words(max_cols) + ENC_FIXED_WORDS;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_slice_encode, (int)(64u * wpg),
                                                     (size_t)wave_words * 4u * wpg) != hipSuccess) {
        (void)hipGetLastError();
        return 0u;
    }
    return (uint32_t)per_cu * tcg_cu_count();
}
