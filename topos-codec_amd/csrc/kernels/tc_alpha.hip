// Alpha plane kernels: MED (JPEG-LS median) prediction inside each band and
// Rice-coded (run, level) pairs with a trailing (run, 0) terminator.
//
// Reference semantics (paths relative to the reference tree):
//   residuals    native/src/codec/codec.c:2969-3013 (fill_alpha_residuals)
//   pair coding  native/src/entropy/block_coding.c:456-561, entropy/rice.h:92-246
//   decode sink  native/src/codec/codec.c:7919-7972 (alpha_pair_sink)
//   mode 2       native/src/codec/codec.c:1598-1617 (pre-quantisation)
#include <hip/hip_runtime.h>

#include "tcg.h"
#include "tc_math.h"
#include "tc_dev.h"

__device__ static inline int32_t med3(int32_t a, int32_t b, int32_t c)
{
    if (c >= a && c >= b) { return a < b ? a : b; }
    if (c <= a && c <= b) { return a < b ? b : a; }
    return a + b - c;
}

__device__ static inline uint32_t k_of(uint64_t sum, uint64_t cnt)
{
    if (cnt == 0u || sum == 0u) { return 0u; }
    const uint64_t mean = sum / cnt;
    uint32_t k = 0u;
    while (k < 14u && (mean >> (k + 1u)) != 0u) { ++k; }
    return k;
}

// Rice code word of m with parameter k: (length, left-aligned 64-bit bits)
__device__ static inline uint32_t rice_word(uint32_t k, uint32_t m, uint64_t* v)
{
    const uint32_t q = m >> k;
    if (q <= 30u) {
        const uint64_t unary = ((uint64_t)((1u << (q + 1u)) - 2u)) << k;
        *v = unary | (uint64_t)(m & ((1u << k) - 1u));
        return q + 1u + k;
    }
    *v = ((uint64_t)0x7FFFFFFFu << 32) | (uint64_t)m;
    return 63u;
}

__device__ static inline uint32_t rice_len(uint32_t k, uint32_t m)
{
    const uint32_t q = m >> k;
    return q <= 30u ? q + 1u + k : 63u;
}

// OR an L-bit code (value in the low L bits of v) into a big-endian bit
// stream held in little-endian 32-bit words, starting at bit position pos.
__device__ static inline void put_code(uint32_t* words, uint64_t pos, uint32_t L, uint64_t v)
{
    if (L == 0u) { return; }
    const uint64_t big = v << (64u - L);
    const uint32_t hi = (uint32_t)(big >> 32), lo = (uint32_t)big;
    const uint64_t w0 = pos >> 5;
    const uint32_t o = (uint32_t)(pos & 31u);
    const uint32_t a = o == 0u ? hi : (hi >> o);
    const uint32_t b = o == 0u ? lo : ((hi << (32u - o)) | (lo >> o));
    const uint32_t c = o == 0u ? 0u : (lo << (32u - o));
    const uint64_t end = pos + L;  // exclusive
    if (a) { atomicOr(&words[w0], bswap32(a)); }
    if (b && ((w0 + 1u) << 5) < end) { atomicOr(&words[w0 + 1u], bswap32(b)); }
    if (c && ((w0 + 2u) << 5) < end) { atomicOr(&words[w0 + 2u], bswap32(c)); }
}

struct AlphaBand {
    const uint16_t* src;
    uint32_t stride, vis_w, vis_h, w, rows, py0, shift;
};

// coded sample: edge replication beyond the visible area, then the mode-2
// pre-quantisation to (16 - shift) significant bits
__device__ static inline int32_t coded_at(const AlphaBand& B, uint32_t x, uint32_t y)
{
    const uint32_t gy = min(B.py0 + y, B.vis_h - 1u);
    const uint32_t gx = min(x, B.vis_w - 1u);
    uint32_t v = B.src[(size_t)gy * B.stride + gx];
    if (B.shift) {
        const uint32_t half = 1u << (B.shift - 1u), top = (0xFFFFu >> B.shift) << B.shift;
        v = ((v + half) >> B.shift) << B.shift;
        if (v > top) { v = top; }
    }
    return (int32_t)v;
}

__device__ static inline int32_t residual_at(const AlphaBand& B, uint32_t i)
{
    const uint32_t y = i / B.w, x = i % B.w;
    int32_t pred;
    if (y == 0u) { pred = x == 0u ? 0 : coded_at(B, x - 1u, 0u); }
    else if (x == 0u) { pred = coded_at(B, 0u, y - 1u); }
    else { pred = med3(coded_at(B, x - 1u, y), coded_at(B, x, y - 1u), coded_at(B, x - 1u, y - 1u)); }
    return coded_at(B, x, y) - pred;
}

__device__ static inline uint32_t raw_src(const AlphaBand& B, uint32_t x, uint32_t y)
{
    const uint32_t gy = min(B.py0 + y, B.vis_h - 1u);
    const uint32_t gx = min(x, B.vis_w - 1u);
    return B.src[(size_t)gy * B.stride + gx];
}

#define AE_T 256

__global__ __launch_bounds__(AE_T) void k_alpha_encode(const tcg_eplane* __restrict__ planes,
                                                        const tcg_eslice* __restrict__ slices,
                                                        uint8_t* __restrict__ scratch,
                                                        tcg_eresult* __restrict__ results)
{
    __shared__ uint64_t sSum[AE_T];
    __shared__ uint32_t sCnt[AE_T];
    __shared__ int32_t sLast[AE_T];
    __shared__ uint64_t sBits[AE_T];
    __shared__ uint32_t sErr[AE_T];
    const tcg_eslice S = slices[blockIdx.x];
    const tcg_eplane P = planes[S.plane];
    if (P.kind != 1u) { return; }
    const uint32_t t = threadIdx.x;
    AlphaBand B;
    B.src = (const uint16_t*)P.src;
    B.stride = P.stride; B.vis_w = P.vis_w; B.vis_h = P.vis_h;
    B.w = P.coded_w; B.rows = S.h * 8u; B.py0 = S.y0 * 8u; B.shift = P.alpha_shift;
    const uint32_t n = B.w * B.rows;
    const uint32_t per = (n + AE_T - 1u) / AE_T;
    const uint32_t a = min(n, t * per), b = min(n, a + per);

    // pass 1: nonzero count, level sum, last nonzero, mode-2 max error
    uint64_t sum = 0u;
    uint32_t cnt = 0u, merr = 0u;
    int32_t last = -1;
    for (uint32_t i = a; i < b; ++i) {
        const int32_t r = residual_at(B, i);
        if (r != 0) { sum += tc_map_signed(r); cnt++; last = (int32_t)i; }
        if (B.shift) {
            const uint32_t y = i / B.w, x = i % B.w;
            const int32_t d = (int32_t)raw_src(B, x, y) - coded_at(B, x, y);
            const uint32_t e = (uint32_t)(d < 0 ? -d : d);
            merr = e > merr ? e : merr;
        }
    }
    sSum[t] = sum; sCnt[t] = cnt; sLast[t] = last; sErr[t] = merr;
    __syncthreads();
    __shared__ uint32_t sK[2];
    __shared__ int32_t sPrevNz[AE_T];
    __shared__ uint32_t sMaxErr;
    if (t == 0u) {
        uint64_t ts = 0u, tc = 0u;
        int32_t pl = -1;
        uint32_t me = 0u;
        for (uint32_t i = 0; i < AE_T; ++i) {
            sPrevNz[i] = pl;
            ts += sSum[i]; tc += sCnt[i];
            if (sLast[i] >= 0) { pl = sLast[i]; }
            me = sErr[i] > me ? sErr[i] : me;
        }
        const uint64_t trailing = (uint64_t)n - (uint64_t)(pl + 1);
        sK[0] = k_of(ts, tc);                                   // k1: levels
        sK[1] = k_of((uint64_t)n - tc, tc + (trailing > 0u ? 1u : 0u));  // k2: runs
        sMaxErr = me;
        sLast[0] = pl;  // global last nonzero
    }
    __syncthreads();
    const uint32_t k1 = sK[0], k2 = sK[1];
    const int32_t glast = sLast[0];
    __syncthreads();
    // pass 2: bits per chunk
    uint64_t bits = 0u;
    {
        int32_t prev = sPrevNz[t];
        for (uint32_t i = a; i < b; ++i) {
            const int32_t r = residual_at(B, i);
            if (r != 0) {
                bits += rice_len(k2, (uint32_t)((int32_t)i - prev - 1)) + rice_len(k1, tc_map_signed(r));
                prev = (int32_t)i;
            }
        }
    }
    sBits[t] = bits;
    __syncthreads();
    __shared__ uint64_t sTotal;
    if (t == 0u) {
        uint64_t acc = 0u;
        for (uint32_t i = 0; i < AE_T; ++i) { const uint64_t v = sBits[i]; sBits[i] = acc; acc += v; }
        const uint64_t trailing = (uint64_t)n - (uint64_t)(glast + 1);
        if (trailing > 0u) { acc += rice_len(k2, (uint32_t)trailing) + rice_len(k1, 0u); }
        sTotal = acc;
    }
    __syncthreads();
    const uint64_t total_bits = sTotal;
    const uint64_t nbytes = (total_bits + 7u) >> 3;
    const uint64_t nwords = (total_bits + 31u) >> 5;
    const bool fits = nwords * 4u + 8u <= S.scratch_cap;
    uint32_t* words = (uint32_t*)(scratch + S.scratch);
    if (fits) {
        for (uint64_t i = t; i < nwords; i += AE_T) { words[i] = 0u; }
    }
    __syncthreads();
    if (fits) {
        uint64_t pos = sBits[t];
        int32_t prev = sPrevNz[t];
        for (uint32_t i = a; i < b; ++i) {
            const int32_t r = residual_at(B, i);
            if (r != 0) {
                uint64_t v;
                uint32_t L = rice_word(k2, (uint32_t)((int32_t)i - prev - 1), &v);
                put_code(words, pos, L, v);
                pos += L;
                L = rice_word(k1, tc_map_signed(r), &v);
                put_code(words, pos, L, v);
                pos += L;
                prev = (int32_t)i;
            }
        }
        if (t == AE_T - 1u) {
            const uint64_t trailing = (uint64_t)n - (uint64_t)(glast + 1);
            if (trailing > 0u) {
                uint64_t p2 = total_bits - rice_len(k2, (uint32_t)trailing) - rice_len(k1, 0u);
                uint64_t v;
                uint32_t L = rice_word(k2, (uint32_t)trailing, &v);
                put_code(words, p2, L, v);
                p2 += L;
                L = rice_word(k1, 0u, &v);
                put_code(words, p2, L, v);
            }
        }
    }
    if (t == 0u) {
        tcg_eresult R;
        R.start = 0u;
        R.size = fits ? (uint32_t)nbytes : 0u;
        R.status = fits ? 0 : -15;
        R.k1 = k1; R.k2 = k2; R.k3 = 0u;
        R.crc = 0u; R.offset = 0u; R.aux = sMaxErr;
        results[blockIdx.x] = R;
    }
}

// ---------------------------------------------------------------------------
// decode: wave-parallel CRC, serial Rice pair parse into a zeroed residual
// band, then MED reconstruction as a skewed wavefront (lane = band row,
// 64 rows per pass; row 0 of a pass takes its upper neighbours from LDS).
// ---------------------------------------------------------------------------
// Bit reader over a payload: a 64-bit cache refilled byte by byte from a
// 256-byte window held one dword per lane (plus the prefetched next window).
struct BitWin {
    uintptr_t base;
    uint32_t nbytes, bytepos;
    uintptr_t wbase;
    uint32_t cur, nxt;
    uint64_t cache;
    uint32_t nb;
    uint64_t consumed;
};

__device__ static inline void bw_init(BitWin& R, const uint8_t* p, uint32_t n)
{
    R.base = (uintptr_t)p;
    R.nbytes = n;
    R.bytepos = 0u;
    R.wbase = R.base & ~(uintptr_t)255u;
    R.cur = *(const uint32_t*)(R.wbase + 4u * __lane_id());
    R.nxt = *(const uint32_t*)(R.wbase + 256u + 4u * __lane_id());
    R.cache = 0u;
    R.nb = 0u;
    R.consumed = 0u;
}

__device__ static inline void bw_refill(BitWin& R)
{
    while (R.nb <= 56u && R.bytepos < R.nbytes) {
        const uintptr_t a = R.base + R.bytepos;
        if (a >= R.wbase + 256u) {
            R.cur = R.nxt;
            R.wbase += 256u;
            R.nxt = *(const uint32_t*)(R.wbase + 256u + 4u * __lane_id());
        }
        const uint32_t off = (uint32_t)(a - R.wbase);
        const uint32_t wd = __builtin_amdgcn_readlane(R.cur, off >> 2);
        const uint64_t byte = (wd >> (8u * (off & 3u))) & 0xFFu;
        R.cache |= byte << (56u - R.nb);
        R.nb += 8u;
        R.bytepos += 1u;
    }
}

__device__ static inline void bw_skip(BitWin& R, uint32_t n)
{
    R.cache = n >= 64u ? 0u : (R.cache << n);
    R.nb -= n;
    R.consumed += n;
}

// Rice decode (q ones, a zero, k bits; 31 ones escape to a 32-bit literal);
// returns 0, -10 (truncated) or -9 (value above m_max)
__device__ static inline int32_t bw_rice(BitWin& R, uint32_t k, uint32_t mmax, uint32_t* m)
{
    uint32_t q = 0u;
    for (;;) {
        bw_refill(R);
        if (R.nb == 0u) { return -10; }
        const uint64_t inv = ~R.cache;
        uint32_t ones = inv == 0u ? 64u : (uint32_t)__clzll(inv);
        if (ones > R.nb) { ones = R.nb; }
        if (q + ones >= 31u) { bw_skip(R, 31u - q); q = 31u; break; }
        if (ones < R.nb) { bw_skip(R, ones + 1u); q += ones; break; }
        bw_skip(R, ones);
        q += ones;
    }
    uint64_t v;
    if (q == 31u) {
        bw_refill(R);
        if (R.nb < 32u) { return -10; }
        v = R.cache >> 32;
        bw_skip(R, 32u);
    } else {
        uint32_t rem = 0u;
        if (k > 0u) {
            bw_refill(R);
            if (R.nb < k) { return -10; }
            rem = (uint32_t)(R.cache >> (64u - k));
            bw_skip(R, k);
        }
        v = ((uint64_t)q << k) | rem;
    }
    if (v > mmax) { return -9; }
    *m = (uint32_t)v;
    return 0;
}

__global__ __launch_bounds__(64) void k_alpha_decode(const TcgConsts* __restrict__ K,
                                                      const tcg_dplane* __restrict__ planes,
                                                      const tcg_dslice* __restrict__ slices,
                                                      int32_t* qbuf,
                                                      int32_t* __restrict__ status)
{
    __shared__ uint32_t sTab[256];
    __shared__ uint32_t sX2n[32];
    __shared__ uint16_t sTop[16384];
    __shared__ int32_t sFail[1];
    const tcg_dslice S = slices[blockIdx.x];
    if (S.kind != 1u) { return; }
    const tcg_dplane P = planes[S.plane];
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 256u; i += 64u) { sTab[i] = K->crc[i]; }
    if (lane < 32u) { sX2n[lane] = K->x2n[lane]; }
    if (lane == 0u) { sFail[0] = 0; }
    __syncthreads();
    const uint8_t* pl = (const uint8_t*)S.payload;
    const uint32_t n = S.size;
    int32_t rc = 0;
    {
        // CRC32, chunked across the wave
        const uint32_t per = (n + 63u) >> 6;
        const uint32_t a = min(n, lane * per), b = min(n, a + per);
        uint32_t c = 0u;
        for (uint32_t i = a; i < b; ++i) { c = sTab[(c ^ pl[i]) & 0xFFu] ^ (c >> 8); }
        uint32_t part = a < b ? tc_crc_shift(sX2n, c, (uint64_t)(n - b)) : 0u;
        for (int off = 32; off > 0; off >>= 1) { part ^= (uint32_t)__shfl_xor((int)part, off); }
        const uint32_t crc = part ^ tc_crc_shift(sX2n, 0xFFFFFFFFu, n) ^ 0xFFFFFFFFu;
        if (crc != S.crc) { rc = -11; }
    }
    const uint32_t w = P.coded_w, rows = S.h * 8u;
    const uint32_t count = w * rows;
    int32_t* res = qbuf + (size_t)S.blk0 * 64u;  // the band's block range holds w * rows ints
    if (rc == 0) {
        for (uint32_t i = lane; i < count; i += 64u) { res[i] = 0; }
        __threadfence_block();
        __syncthreads();
        BitWin R;
        bw_init(R, pl, n);
        uint32_t pos = 0u;
        while (pos < count) {
            const uint32_t remaining = count - pos;
            uint32_t rm = 0u, lm = 0u;
            int32_t e = bw_rice(R, S.k2, remaining, &rm);
            if (e == 0) { e = bw_rice(R, S.k1, 0x00020000u, &lm); }
            if (e != 0) { rc = e; break; }
            if (lm == 0u) {
                if (rm != remaining) { rc = -9; }
                break;
            }
            if ((uint64_t)rm + 1u > remaining) { rc = -9; break; }
            if (lane == 0u) { res[pos + rm] = tc_unmap_signed(lm); }
            pos += rm + 1u;
        }
        if (rc == 0) {
            const uint64_t aligned = (R.consumed + 7u) & ~(uint64_t)7u;
            if (aligned > (uint64_t)n * 8u) { rc = -10; }
            else if (aligned != (uint64_t)n * 8u) { rc = -9; }
        }
        __threadfence();
        __syncthreads();
    }
    if (rc == 0) {
        uint16_t* dst = (uint16_t*)P.dst;
        const uint32_t py0 = S.y0 * 8u;
        const uint32_t mask = P.alpha_mask;
        for (uint32_t g = 0; g < rows; g += 64u) {
            const uint32_t y = g + lane;
            const bool row_ok = y < rows;
            int32_t out1 = 0, out2 = 0;  // this lane's values at steps t-1, t-2
            uint32_t bad = 0u;
            const uint32_t steps = w + 64u;
            for (uint32_t t = 0; t < steps; ++t) {
                const int32_t x = (int32_t)t - (int32_t)lane;
                const int32_t up1 = __shfl_up(out1, 1);
                const int32_t up2 = __shfl_up(out2, 1);
                int32_t v = out1;
                if (row_ok && x >= 0 && x < (int32_t)w) {
                    int32_t top, topleft;
                    if (lane == 0u) {
                        top = g > 0u ? (int32_t)sTop[x] : 0;
                        topleft = (g > 0u && x > 0) ? (int32_t)sTop[x - 1] : 0;
                    } else {
                        top = up1;
                        topleft = up2;
                    }
                    const int32_t left = out1;
                    int32_t pred;
                    if (y == 0u) { pred = x == 0 ? 0 : left; }
                    else if (x == 0) { pred = top; }
                    else { pred = med3(left, top, topleft); }
                    v = pred + res[(size_t)y * w + (uint32_t)x];
                    v = v < 0 ? 0 : (v > 65535 ? 65535 : v);
                    if (mask != 0u && ((uint32_t)v & mask) != 0u) { bad = 1u; }
                    const uint32_t gy = py0 + y;
                    if (gy < P.vis_h && (uint32_t)x < P.vis_w) { dst[(size_t)gy * P.stride + (uint32_t)x] = (uint16_t)v; }
                    if (lane == 63u) { sTop[x] = (uint16_t)v; }
                }
                out2 = out1;
                out1 = v;
            }
            __syncthreads();
            if (__ballot(bad != 0u) != 0ull) { rc = -9; break; }
        }
    }
    if (rc != 0) {
        // concealment: the band's visible rows take the fill value (0xFFFF),
        // over anything a failed reconstruction already stored
        // (native/src/codec/codec.c:7975-7985)
        uint16_t* dst = (uint16_t*)P.dst;
        const uint32_t py0 = S.y0 * 8u;
        const uint32_t nrows = py0 < P.vis_h ? min(rows, P.vis_h - py0) : 0u;
        for (uint32_t y = 0; y < nrows; ++y) {
            for (uint32_t x = lane; x < P.vis_w; x += 64u) { dst[(size_t)(py0 + y) * P.stride + x] = (uint16_t)P.fill; }
        }
    }
    if (lane == 0u) { status[blockIdx.x] = rc == 0 ? 0 : 1; }
}

void tcg_alpha_encode_launch(hipStream_t st, const tcg_eplane* planes, const tcg_eslice* slices,
                             uint32_t nslices, const int32_t* unused, uint8_t* scratch,
                             tcg_eresult* results)
{
    (void)unused;
    hipLaunchKernelGGL(k_alpha_encode, dim3(nslices), dim3(AE_T), 0, st, planes, slices, scratch, results);
}

void tcg_alpha_decode_launch(hipStream_t st, const tcg_dplane* planes, const tcg_dslice* slices,
                             uint32_t nslices, int32_t* qbuf, int32_t* status)
{
    hipLaunchKernelGGL(k_alpha_decode, dim3(nslices), dim3(64), 0, st, tcg_consts_dev(), planes, slices,
                       qbuf, status);
}
