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
// wave-wide CRC32 of a byte range: each lane folds a contiguous chunk with the
// bytewise table, the partial registers are shifted over the bytes that follow
// them (multiplication by x^(8k) mod P) and XOR-reduced across the wave.
// ---------------------------------------------------------------------------
__device__ static uint32_t wave_crc32(const uint8_t* p, uint32_t n, const uint32_t* tab,
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

// ---------------------------------------------------------------------------
// K5: rANS2 colour slice decoder. The 32-bit state is wave-uniform; each lane
// holds one symbol's cumulative frequency so a slot is resolved with a single
// compare + ballot. The payload is read through a 256-byte window held one
// dword per lane (plus a prefetched next window). Coefficients are assembled
// in registers (lane = natural index) and written as one 256-byte block.
// ---------------------------------------------------------------------------
struct DecStream {
    uint32_t x;
    uint32_t pos;     // next byte index inside the payload
    uint32_t n;       // payload bytes
    uintptr_t base;   // payload address
    uintptr_t wbase;  // 256-aligned address of the current window
    uint32_t cur, nxt;
    int32_t err;
};

__device__ static inline void ds_window_load(DecStream& D, uintptr_t w)
{
    D.wbase = w;
    D.cur = *(const uint32_t*)(w + 4u * __lane_id());
    D.nxt = *(const uint32_t*)(w + 256u + 4u * __lane_id());
}

__device__ static inline uint32_t ds_byte(DecStream& D)
{
    const uintptr_t a = D.base + D.pos;
    if (a >= D.wbase + 256u) {
        D.cur = D.nxt;
        D.wbase += 256u;
        D.nxt = *(const uint32_t*)(D.wbase + 256u + 4u * __lane_id());
    }
    const uint32_t off = (uint32_t)(a - D.wbase);
    const uint32_t w = __builtin_amdgcn_readlane(D.cur, off >> 2);
    D.pos += 1u;
    return (w >> (8u * (off & 3u))) & 0xFFu;
}

__device__ static inline void ds_renorm(DecStream& D)
{
    while (D.x < (1u << 23)) {
        if (D.pos >= D.n) { D.err = -10; return; }
        D.x = (D.x << 8) | ds_byte(D);
    }
}

__device__ static inline uint32_t ds_sym(DecStream& D, const uint16_t* freq, const uint16_t* cum,
                                         uint32_t nsym)
{
    const uint32_t lane = __lane_id();
    const uint32_t slot = D.x & 4095u;
    const uint32_t cv = lane < nsym ? (uint32_t)cum[lane] : 0xFFFFu;
    const uint64_t m = __ballot(cv <= slot);
    const uint32_t s = 63u - (uint32_t)__clzll(m);
    const uint32_t f = freq[s], c = cum[s];
    D.x = f * (D.x >> 12) + slot - c;
    ds_renorm(D);
    return s;
}

__device__ static inline uint32_t ds_chunk(DecStream& D, uint32_t w)
{
    const uint32_t c = D.x & ((1u << w) - 1u);
    D.x >>= w;
    ds_renorm(D);
    return c;
}

__device__ static inline uint32_t ds_suffix(DecStream& D, uint32_t cat)
{
    uint32_t v = 0u;
    if (cat > 1u) {
        const uint32_t b = cat - 1u;
        if (b > 23u) {
            const uint32_t lo = ds_chunk(D, 16u);
            if (D.err) { return 0u; }
            const uint32_t hi = ds_chunk(D, b - 16u);
            v = (hi << 16) | lo;
        } else {
            v = ds_chunk(D, b);
        }
    }
    return cat == 0u ? 0u : ((1u << (cat - 1u)) | v);
}

__global__ __launch_bounds__(64) void k_slice_decode(const TcgConsts* __restrict__ K,
                                                      const tcg_dplane* __restrict__ planes,
                                                      const tcg_dslice* __restrict__ slices,
                                                      int32_t* __restrict__ qbuf,
                                                      int32_t* __restrict__ status)
{
    __shared__ uint32_t sTab[256];
    __shared__ uint32_t sX2n[32];
    __shared__ uint16_t sFreq[11][64];
    __shared__ uint16_t sCum[11][65];
    __shared__ int32_t sDcRow[2048];
    __shared__ int32_t sBad[1];

    const uint32_t lane = threadIdx.x;
    const tcg_dslice S = slices[blockIdx.x];
    if (S.kind != 0u) { return; }
    const tcg_dplane P = planes[S.plane];
    for (uint32_t i = lane; i < 256u; i += 64u) { sTab[i] = K->crc[i]; }
    if (lane < 32u) { sX2n[lane] = K->x2n[lane]; }
    if (lane == 0u) { sBad[0] = 0; }
    __syncthreads();

    const uint8_t* pl = (const uint8_t*)S.payload;
    const uint32_t n = S.size;
    int32_t rc = 0;
    if (wave_crc32(pl, n, sTab, sX2n) != S.crc) { rc = -11; }
    uint32_t lvl_model = 0u, dc_model = 0u, lvl_n = 1u, dc_n = 1u, prefix = 0u;
    if (rc == 0) {
        if (n < 1u + 121u + 4u) { rc = -10; }
    }
    if (rc == 0) {
        const uint32_t flags = pl[0];
        lvl_model = flags & 3u;
        dc_model = (flags >> 2) & 1u;
        if ((flags & 0xF8u) != 0u || lvl_model > 2u) { rc = -9; }
        lvl_n = lvl_model == 1u ? 4u : (lvl_model == 2u ? 5u : 1u);
        dc_n = dc_model ? 5u : 1u;
        prefix = 1u + dc_n * 29u + 64u + lvl_n * 28u;
        if (rc == 0 && n < prefix + 4u) { rc = -10; }
    }
    if (rc == 0) {
        // rows: [0, dc_n) dc, dc_n run, then lvl rows; lane r builds row r
        const uint32_t rows = dc_n + 1u + lvl_n;
        if (lane < rows) {
            uint32_t off, nsym;
            if (lane < dc_n) { off = 1u + lane * 29u; nsym = 29u; }
            else if (lane == dc_n) { off = 1u + dc_n * 29u; nsym = 64u; }
            else { off = 1u + dc_n * 29u + 64u + (lane - dc_n - 1u) * 28u; nsym = 28u; }
            uint8_t rb[64];
            for (uint32_t s = 0; s < nsym; ++s) { rb[s] = pl[off + s]; }
            if (tc_model_from_bytes(rb, nsym, sFreq[lane], sCum[lane]) != 0) { atomicOr(&sBad[0], 1); }
        }
        __syncthreads();
        if (sBad[0] != 0) { rc = -9; }
    }
    DecStream D;
    D.err = 0;
    if (rc == 0) {
        D.base = (uintptr_t)(pl + prefix);
        D.n = n - prefix;
        D.x = ((uint32_t)pl[prefix] << 24) | ((uint32_t)pl[prefix + 1u] << 16) |
              ((uint32_t)pl[prefix + 2u] << 8) | (uint32_t)pl[prefix + 3u];
        D.pos = 4u;
        if (D.x < (1u << 23)) { rc = -9; }
        ds_window_load(D, (D.base + D.pos) & ~(uintptr_t)255u);
    }
    if (rc == 0) {
        const uint32_t cols = P.bcols;
        const uint32_t dc_max = S.dc_max;
        const uint32_t lv_max = S.lv_max;
        const int64_t dlim = dc_max > 0x08000000u ? ((int64_t)1 << 27) : ((int64_t)1 << 25);
        const uint32_t run_row = dc_n, lvl_row0 = dc_n + 1u;
        uint32_t pdc = 0u;
        int32_t left = 0;
        for (uint32_t by = 0; by < S.h && rc == 0; ++by) {
            for (uint32_t bx = 0; bx < cols; ++bx) {
                const uint32_t dctx = (by == 0u && bx == 0u) ? 0u : tc_ctx_dc(pdc, 1u);
                const uint32_t drow = dc_model ? dctx : 0u;
                uint32_t cat = ds_sym(D, sFreq[drow], sCum[drow], 29u);
                if (D.err) { rc = D.err; break; }
                uint32_t m = ds_suffix(D, cat);
                if (D.err) { rc = D.err; break; }
                if (m > dc_max) { rc = -9; break; }
                const int32_t top = by ? sDcRow[bx] : 0;
                const int32_t pred = tc_dc_predict(bx != 0u, left, by != 0u, top);
                const int64_t dv = (int64_t)pred + (int64_t)tc_unmap_signed(m);
                if (dv > dlim || dv < -dlim) { rc = -9; break; }
                const int32_t dc = (int32_t)dv;
                int32_t val = lane == 0u ? dc : 0;
                __syncthreads();
                if (lane == 0u) { sDcRow[bx] = dc; }
                left = dc;
                pdc = cat;
                uint32_t pos = 1u, pcat = 0u, hp = 0u;
                for (;;) {
                    const uint32_t rs = ds_sym(D, sFreq[run_row], sCum[run_row], 64u);
                    if (D.err) { rc = D.err; break; }
                    if (rs == 63u) { break; }
                    const uint32_t idx = pos + rs;
                    if (idx > 63u) { rc = -9; break; }
                    uint32_t lc = 0u;
                    if (lvl_model == 1u) { lc = tc_ctx_pos(idx); }
                    else if (lvl_model == 2u) { lc = tc_ctx_prev(pcat, hp); }
                    cat = ds_sym(D, sFreq[lvl_row0 + lc], sCum[lvl_row0 + lc], 28u);
                    if (D.err) { rc = D.err; break; }
                    m = ds_suffix(D, cat);
                    if (D.err) { rc = D.err; break; }
                    if (m > lv_max) { rc = -9; break; }
                    if (lane == (uint32_t)kZigzag[idx]) { val = tc_unmap_signed(m); }
                    pcat = cat;
                    hp = 1u;
                    pos = idx + 1u;
                }
                if (rc != 0) { break; }
                qbuf[((size_t)S.blk0 + (size_t)by * cols + bx) * 64u + lane] = val;
            }
        }
        if (rc == 0 && D.pos != D.n) { rc = -9; }
    }
    if (lane == 0u) { status[blockIdx.x] = rc == 0 ? 0 : 1; }
}

// ---------------------------------------------------------------------------
// K6: dequantisation + inverse transform + level shift + clamp + clip-store.
// Tile = 32 blocks of one block row, 256 threads. Pass 1: thread (j, u)
// dequantises row u of block j and forms B[u][x] = sum_v G[u][v] M[v][x];
// pass 2: thread (j, x) forms x[y][x] = sum_u M[u][y] B[u][x] and writes
// column x of the block. For bit depth <= 12 every intermediate is an integer
// below 2^53 in magnitude, so f64 FMA is exact; deeper samples use int64.
// ---------------------------------------------------------------------------
#define IQ_BLOCKS 32

__device__ static inline uint32_t find_dplane(const tcg_dplane* planes, uint32_t nplanes, uint32_t tile)
{
    uint32_t lo = 0, hi = nplanes - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (planes[mid].tile0 <= tile) { lo = mid; } else { hi = mid - 1; }
    }
    return lo;
}

template <bool WIDE>
__device__ static void dequant_idct_tile(const tcg_dplane& P, const tcg_qctx* __restrict__ qc,
                                         uint32_t brow, uint32_t bx0, uint32_t nb, bool conceal,
                                         const int32_t* __restrict__ qbuf, void* sBraw, uint32_t* sQ)
{
    typedef typename std::conditional<WIDE, int64_t, double>::type acc_t;
    acc_t* sB = (acc_t*)sBraw;
    const uint32_t t = threadIdx.x;
    if (t < 64u) { sQ[t] = qc->Q[t]; }
    const int32_t mid = qc->mid, maxv = qc->maxv;
    const int64_t fclamp = qc->fclamp;
    uint16_t* out = (uint16_t*)(P.deblock ? P.tmp : P.dst);
    const uint32_t y0 = brow * 8u;
    if (conceal) {
        // fill the visible part of the tile with the neutral value
        const uint32_t xs = bx0 * 8u, xe = min(P.vis_w, (bx0 + nb) * 8u);
        for (uint32_t i = t; i < 8u * 256u; i += 256u) {
            const uint32_t y = y0 + i / 256u, x = xs + i % 256u;
            if (y < P.vis_h && x < xe) { out[(size_t)y * P.stride + x] = (uint16_t)P.fill; }
        }
        return;
    }
    __syncthreads();
    {
        const uint32_t j = t >> 3, u = t & 7u;
        if (j < nb) {
            const int32_t* src = qbuf + ((size_t)P.blk0 + (size_t)brow * P.bcols + bx0 + j) * 64u + u * 8u;
            const int4 r0 = *(const int4*)src;
            const int4 r1 = *(const int4*)(src + 4);
            const int32_t qv[8] = { r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w };
            acc_t G[8];
#pragma unroll
            for (int v = 0; v < 8; ++v) {
                int64_t d = (int64_t)qv[v] * (int64_t)sQ[u * 8 + v];
                d = d > fclamp ? fclamp : (d < -fclamp ? -fclamp : d);
                if (WIDE) {
                    d = d > (1 << 30) ? (1 << 30) : (d < -(1 << 30) ? -(1 << 30) : d);
                }
                G[v] = (acc_t)d * (acc_t)kTW[u * 8 + v];
            }
#pragma unroll
            for (int x = 0; x < 8; ++x) {
                acc_t acc = 0;
#pragma unroll
                for (int v = 0; v < 8; ++v) { acc += G[v] * (acc_t)kTM[v * 8 + x]; }
                sB[j * 65 + u * 8 + x] = acc;
            }
        }
    }
    __syncthreads();
    {
        const uint32_t j = t >> 3, x = t & 7u;
        if (j < nb) {
            acc_t B[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) { B[u] = sB[j * 65 + u * 8 + x]; }
            const uint32_t px = (bx0 + j) * 8u + x;
#pragma unroll
            for (int y = 0; y < 8; ++y) {
                acc_t acc = 0;
#pragma unroll
                for (int u = 0; u < 8; ++u) { acc += (acc_t)kTM[u * 8 + y] * B[u]; }
                int32_t r;
                if (WIDE) {
                    const int64_t v = (int64_t)acc;
                    r = v >= 0 ? (int32_t)((v + ((int64_t)1 << 31)) >> 32)
                               : -(int32_t)((-v + ((int64_t)1 << 31)) >> 32);
                } else {
                    const double v = (double)acc;
                    r = v >= 0.0 ? (int32_t)floor((v + 2147483648.0) * (1.0 / 4294967296.0))
                                 : -(int32_t)floor((-v + 2147483648.0) * (1.0 / 4294967296.0));
                }
                int32_t s = r + mid;
                s = s < 0 ? 0 : (s > maxv ? maxv : s);
                const uint32_t py = y0 + (uint32_t)y;
                if (py < P.vis_h && px < P.vis_w) { out[(size_t)py * P.stride + px] = (uint16_t)s; }
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_dequant_idct(const tcg_qctx* __restrict__ qctxs,
                                                       const tcg_dplane* __restrict__ planes,
                                                       uint32_t nplanes,
                                                       const tcg_dslice* __restrict__ slices,
                                                       const uint16_t* __restrict__ row_slice,
                                                       const int32_t* __restrict__ status,
                                                       const int32_t* __restrict__ qbuf)
{
    __shared__ uint64_t sB[IQ_BLOCKS * 65];
    __shared__ uint32_t sQ[64];
    const uint32_t tile = blockIdx.x;
    const uint32_t pi = find_dplane(planes, nplanes, tile);
    const tcg_dplane P = planes[pi];
    if (P.kind != 0u) { return; }
    const uint32_t tiles_per_row = (P.bcols + IQ_BLOCKS - 1u) / IQ_BLOCKS;
    const uint32_t local = tile - P.tile0;
    const uint32_t brow = local / tiles_per_row;
    const uint32_t bx0 = (local % tiles_per_row) * IQ_BLOCKS;
    const uint32_t nb = min((uint32_t)IQ_BLOCKS, P.bcols - bx0);
    const uint32_t si = row_slice[P.row_slice + brow];
    const tcg_qctx* qc = &qctxs[slices[si].qctx];
    const bool conceal = status[si] != 0;
    if (qc->bit_depth > 12u) {
        dequant_idct_tile<true>(P, qc, brow, bx0, nb, conceal, qbuf, sB, sQ);
    } else {
        dequant_idct_tile<false>(P, qc, brow, bx0, nb, conceal, qbuf, sB, sQ);
    }
}

// ---------------------------------------------------------------------------
// K7: output deblocking, vertical edges then horizontal edges, fused per
// strip. Strip r covers rows [8r-2, 8r+6): every horizontal edge y0 = 8r
// reads and writes rows y0-2..y0+1 only, so the strips are independent once
// each has applied the vertical-edge filter to its own rows. A tile is one
// strip x 256 columns plus a 2-column halo on each side; reads come from the
// unfiltered plane and writes go to the output plane.
// ---------------------------------------------------------------------------
#define DB_COLS 256

__device__ static inline uint32_t db_absu(int32_t v) { return (uint32_t)(v < 0 ? -v : v); }
__device__ static inline int32_t db_clampv(int32_t v, int32_t m) { return v < 0 ? 0 : (v > m ? m : v); }

__device__ static inline void db_edge(int32_t& p1, int32_t& p0, int32_t& q0, int32_t& q1, int32_t alpha,
                                      int32_t beta, int32_t clip, int32_t maxv)
{
    if (!(db_absu(q0 - p0) < (uint32_t)alpha && db_absu(p1 - p0) < (uint32_t)beta &&
          db_absu(q1 - q0) < (uint32_t)beta)) {
        return;
    }
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

__global__ __launch_bounds__(256) void k_deblock(const tcg_dplane* __restrict__ planes, uint32_t nplanes,
                                                  const uint32_t* __restrict__ tile0)
{
    __shared__ int32_t sP[8][DB_COLS + 4];
    const uint32_t tile = blockIdx.x;
    uint32_t pi = 0;
    {
        uint32_t lo = 0, hi = nplanes - 1;
        while (lo < hi) {
            const uint32_t mid = (lo + hi + 1) >> 1;
            if (tile0[mid] <= tile) { lo = mid; } else { hi = mid - 1; }
        }
        pi = lo;
    }
    const tcg_dplane P = planes[pi];
    if (!P.deblock) { return; }
    const uint32_t w = P.vis_w, h = P.vis_h;
    const uint32_t col_tiles = (w + DB_COLS - 1u) / DB_COLS;
    const uint32_t local = tile - tile0[pi];
    const uint32_t r = local / col_tiles;
    const uint32_t c0 = (local % col_tiles) * DB_COLS;
    const int32_t ys = (int32_t)(8u * r) - 2 < 0 ? 0 : (int32_t)(8u * r) - 2;
    const int32_t ye = min((int32_t)h, (int32_t)(8u * r) + 6);
    const int32_t nrows = ye - ys;
    const int32_t xs = (int32_t)c0 - 2;
    const uint16_t* src = (const uint16_t*)P.tmp;
    uint16_t* dst = (uint16_t*)P.dst;
    const int32_t alpha = (int32_t)P.db_alpha, beta = alpha / 2 + 2, clip = alpha >> 1;
    const int32_t maxv = (int32_t)P.maxv;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < 8u * (DB_COLS + 4u); i += 256u) {
        const int32_t rr = (int32_t)(i / (DB_COLS + 4u)), cc = (int32_t)(i % (DB_COLS + 4u));
        const int32_t y = ys + rr, x = xs + cc;
        sP[rr][cc] = (rr < nrows && x >= 0 && x < (int32_t)w) ? (int32_t)src[(size_t)y * P.stride + x] : 0;
    }
    __syncthreads();
    // vertical edges x0 = c0, c0+8, ..., c0+256 (x0 >= 8, x0 + 1 < w), all rows
    for (uint32_t i = t; i < 8u * 33u; i += 256u) {
        const int32_t rr = (int32_t)(i / 33u);
        const int32_t x0 = (int32_t)c0 + 8 * (int32_t)(i % 33u);
        if (rr < nrows && x0 >= 8 && x0 + 1 < (int32_t)w) {
            const int32_t l = x0 - xs;
            int32_t p1 = sP[rr][l - 2], p0 = sP[rr][l - 1], q0 = sP[rr][l], q1 = sP[rr][l + 1];
            db_edge(p1, p0, q0, q1, alpha, beta, clip, maxv);
            sP[rr][l - 2] = p1; sP[rr][l - 1] = p0; sP[rr][l] = q0; sP[rr][l + 1] = q1;
        }
    }
    __syncthreads();
    // horizontal edge y0 = 8r for 1 <= r <= h/8 - 1
    if (r >= 1u && r + 1u <= h / 8u) {
        const int32_t x = (int32_t)c0 + (int32_t)t;
        if (x < (int32_t)w) {
            const int32_t l = x - xs;
            // local rows of y0-2 .. y0+1 are 0..3 (ys = y0 - 2)
            int32_t p1 = sP[0][l], p0 = sP[1][l], q0 = sP[2][l], q1 = sP[3][l];
            db_edge(p1, p0, q0, q1, alpha, beta, clip, maxv);
            sP[0][l] = p1; sP[1][l] = p0; sP[2][l] = q0; sP[3][l] = q1;
        }
    }
    __syncthreads();
    for (uint32_t i = t; i < 8u * DB_COLS; i += 256u) {
        const int32_t rr = (int32_t)(i / DB_COLS), x = (int32_t)c0 + (int32_t)(i % DB_COLS);
        if (rr < nrows && x < (int32_t)w) {
            dst[(size_t)(ys + rr) * P.stride + x] = (uint16_t)sP[rr][x - xs];
        }
    }
}

// ---------------------------------------------------------------------------
// host launch
// ---------------------------------------------------------------------------
extern "C" int32_t tcg_decode(const tcg_qctx* qctx_dev, uint32_t nplanes, const tcg_dplane* planes_dev,
                              uint32_t nslices, const tcg_dslice* slices_dev,
                              const uint16_t* row_slice_dev, uint64_t qbuf, uint32_t total_tiles,
                              int32_t* slice_status_dev, uint32_t db_tiles, const uint32_t* db_tile0_dev)
{
    hipStream_t st = tcg_stream();
    const TcgConsts* K = tcg_consts_dev();
    TcgTimer tm;
    tm.mark(st, 0);
    hipLaunchKernelGGL(k_slice_decode, dim3(nslices), dim3(64), 0, st, K, planes_dev, slices_dev,
                       (int32_t*)qbuf, slice_status_dev);
    tcg_alpha_decode_launch(st, planes_dev, slices_dev, nslices, (int32_t*)qbuf, slice_status_dev);
    tm.mark(st, 1);
    if (total_tiles > 0u) {
        hipLaunchKernelGGL(k_dequant_idct, dim3(total_tiles), dim3(256), 0, st, qctx_dev, planes_dev,
                           nplanes, slices_dev, row_slice_dev, slice_status_dev, (const int32_t*)qbuf);
    }
    tm.mark(st, 2);
    if (db_tiles > 0u) {
        hipLaunchKernelGGL(k_deblock, dim3(db_tiles), dim3(256), 0, st, planes_dev, nplanes, db_tile0_dev);
    }
    tm.mark(st, 3);
    tm.finish_decode();
    return tcg_check("decode launch");
}
