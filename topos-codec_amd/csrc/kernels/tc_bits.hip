// gfx950 kernels for the bit-packed colour entropy modes: V1 bounded Rice
// (major 1) and V2 canonical VLC (major 2, entropy mode 1), one wavefront per
// slice in both directions. Transform, quantisation, reconstruction, alpha
// and packet assembly are the kernels the rANS2 path uses.
//
// Reference semantics (paths relative to the reference tree):
//   symbols        docs/bitstream_spec_v1.md:286-305, native/src/codec/codec.c:1702-1848
//   Rice           native/src/entropy/rice.h:92-116, codec.c:734-741,1854-1881
//   VLC            docs/bitstream_spec_v2.md:64-110, native/src/entropy/vlc.c:37-151,
//                  codec.c:1896-2015
//   slice decode   native/src/bitstream/slice_codec.c:233-341,
//                  native/src/entropy/block_coding.c:78-111,146-184
#include <hip/hip_runtime.h>

#include "tcg.h"
#include "tc_math.h"
#include "tc_dev.h"

// ---------------------------------------------------------------------------
// K2b: V1 / V2 slice encoder, one wavefront per slice, lane i = scan position
// i of the current block.
// Forward pass: DC prediction and the statistics the mode needs (Rice:
// per-family sums for k = floor(log2(mean)); VLC: symbol histograms for the
// exact book choice). Emit pass: every lane forms the code words of its own
// symbols (lane 0 the DC residual, a lane holding a nonzero coefficient its
// run and level), a wave prefix sum of the word lengths gives each word its
// bit position, and the words are OR-ed into a ring of 32-bit words in LDS
// (MSB-first). Whole 64-word groups leave as one coalesced 256-byte store.
// ---------------------------------------------------------------------------
#define BW_RING 1024u  // words; one block writes at most 253 words
#define BW_MASK (BW_RING - 1u)
#define BE_GROUP 8u

// Rice code word of m: q = m >> k ones, a zero and the k low bits of m; q > 30
// escapes to 31 ones and the 32-bit literal (native/src/entropy/rice.h:104-116)
__device__ static inline uint32_t rice_word(uint32_t k, uint32_t m, uint64_t& v)
{
    const uint32_t q = m >> k;
    if (q <= 30u) {
        v = ((((uint64_t)1 << (q + 1u)) - 2u) << k) | (uint64_t)(m & ((1u << k) - 1u));
        return q + 1u + k;
    }
    v = ((uint64_t)0x7FFFFFFFu << 32) | m;
    return 63u;
}

// category word of m under a VLC book (code of cat = bitlen(m), then the cat - 1
// low bits of m)
__device__ static inline uint32_t vlc_cat_word(const uint32_t* code, const uint32_t* len, uint32_t m, uint64_t& v)
{
    const uint32_t cat = tc_bitlen(m);
    const uint32_t sb = cat > 1u ? cat - 1u : 0u;
    v = ((uint64_t)code[cat] << sb) | (uint64_t)(m & ((1u << sb) - 1u));
    return len[cat] + sb;
}

// OR an n-bit word (1 <= n <= 64, right-aligned in v) into the ring at bit b
__device__ static inline void ring_put(uint32_t* ring, uint64_t b, uint32_t n, uint64_t v)
{
    const uint64_t hi = v << (64u - n);
    const uint32_t sh = (uint32_t)(b & 31u);
    const uint32_t w = (uint32_t)(b >> 5);
    const uint64_t a = hi >> sh;
    atomicOr(&ring[w & BW_MASK], (uint32_t)(a >> 32));
    if ((uint32_t)a != 0u) { atomicOr(&ring[(w + 1u) & BW_MASK], (uint32_t)a); }
    if (sh != 0u) {
        const uint32_t t = (uint32_t)((hi << (64u - sh)) >> 32);
        if (t != 0u) { atomicOr(&ring[(w + 2u) & BW_MASK], t); }
    }
}

__device__ static inline uint64_t wave_sum64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ static inline uint64_t readlane64(uint64_t v, uint32_t l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// floor(log2(sum / count)) capped at 14; 0 for an empty family
// (native/src/codec/codec.c:734-741)
__device__ static inline uint32_t rice_k(uint64_t sum, uint64_t count)
{
    if (count == 0u || sum == 0u) { return 0u; }
    const uint64_t mean = sum / count;
    if (mean < 2u) { return 0u; }
    const uint32_t l = 63u - (uint32_t)__builtin_clzll(mean);
    return l < 14u ? l : 14u;
}

// histogram layout (u32): DC category 29, run 64, level category 28
#define BH_DC 0
#define BH_RUN 29
#define BH_LVL 93
#define BH_TOTAL 121

__global__ __launch_bounds__(64) void k_slice_encode_bits(const TcgConsts* __restrict__ K,
                                                           const tcg_eplane* __restrict__ planes,
                                                           const tcg_eslice* __restrict__ slices,
                                                           const int32_t* __restrict__ qbuf,
                                                           uint32_t* __restrict__ dcmbuf,
                                                           uint8_t* __restrict__ scratch,
                                                           tcg_eresult* __restrict__ results)
{
    __shared__ uint32_t sRing[BW_RING];
    __shared__ uint32_t sH[BH_TOTAL];
    extern __shared__ int32_t sDcRow[];  // one per block column of the batch's widest plane
    __shared__ uint32_t sCode[3][64], sLen[3][64];  // chosen books: 0 DC, 1 run, 2 level

    const uint32_t lane = threadIdx.x;
    const tcg_eslice S = slices[blockIdx.x];
    if (S.mode == TCG_MODE_RANS2) { return; }
    const tcg_eplane P = planes[S.plane];
    if (P.kind != 0u) { return; }
    const uint32_t vlc = S.mode == TCG_MODE_VLC ? 1u : 0u;
    const uint32_t cols = P.bcols;
    const uint32_t nblk = S.nblk;
    for (uint32_t i = lane; i < BH_TOTAL; i += 64u) { sH[i] = 0u; }
    for (uint32_t i = lane; i < BW_RING; i += 64u) { sRing[i] = 0u; }
    __syncthreads();

    const uint64_t below = lane == 0u ? 0ull : ((1ull << lane) - 1ull);
    const int32_t* qs = qbuf + (size_t)S.blk0 * 64u + lane;

    // ---------------- forward pass ----------------
    uint32_t viol = 0u;
    uint64_t s_dc = 0u, s_lv = 0u, s_run = 0u;
    uint32_t n_pair = 0u;
    int32_t left = 0;
    uint32_t bx = 0u, by = 0u;
    auto fwd_block = [&](int32_t q, uint32_t b) {
        const int32_t q0 = __builtin_amdgcn_readlane(q, 0);
        const int32_t top = by ? (int32_t)wuni((uint32_t)sDcRow[bx]) : 0;
        const int32_t pred = tc_dc_predict(bx != 0u, left, by != 0u, top);
        uint32_t dcm = tc_map_signed(q0 - pred);
        if (dcm > 0x0FFFFFFFu) { dcm = 0x0FFFFFFFu; viol = 1u; }
        if (lane == 0u) {
            dcmbuf[S.blk0 + b] = dcm;
            sDcRow[bx] = q0;
        }
        left = q0;
        const uint64_t nz = __ballot(q != 0) & ~1ull;
        const uint32_t lm_raw = tc_map_signed(q);
        const uint32_t lm = min(lm_raw, 0x07FFFFFFu);
        const uint64_t bl = nz & below;
        const uint32_t pp = bl ? 63u - (uint32_t)__clzll(bl) : 0u;
        const uint32_t run = lane - pp - 1u;
        const bool active = lane != 0u && q != 0;
        if (active && lm_raw > 0x07FFFFFFu) { viol = 1u; }
        if (vlc) {
            if (lane == 0u) {
                atomicAdd(&sH[BH_DC + tc_bitlen(dcm)], 1u);
                atomicAdd(&sH[BH_RUN + 63u], 1u);
            }
            if (active) {
                atomicAdd(&sH[BH_RUN + run], 1u);
                atomicAdd(&sH[BH_LVL + tc_bitlen(lm)], 1u);
            }
        } else {
            s_dc += dcm;
            if (active) {
                s_lv += lm;
                s_run += run;
                n_pair += 1u;
            }
        }
        if (++bx == cols) { bx = 0u; ++by; }
    };
    {
        int32_t qa[BE_GROUP], qb[BE_GROUP];
#pragma unroll
        for (uint32_t j = 0; j < BE_GROUP; ++j) { qa[j] = j < nblk ? qs[(size_t)j * 64u] : 0; }
        for (uint32_t g = 0; g < nblk; g += BE_GROUP) {
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) {
                const uint32_t b = g + BE_GROUP + j;
                qb[j] = b < nblk ? qs[(size_t)b * 64u] : 0;
            }
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) {
                if (g + j < nblk) { fwd_block(qa[j], g + j); }
            }
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) { qa[j] = qb[j]; }
        }
    }
    __syncthreads();
    const uint32_t any_viol = __ballot(viol != 0u) != 0ull ? 1u : 0u;

    // ---------------- parameters ----------------
    uint32_t k1, k2, k3;
    if (!vlc) {
        const uint64_t lv = wave_sum64(s_lv);
        const uint64_t rn = wave_sum64(s_run) + 63ull * nblk;
        const uint64_t np = wave_sum64(n_pair);
        k1 = rice_k(s_dc, nblk);
        k2 = rice_k(lv, np);
        k3 = rice_k(rn, np + nblk);
    } else {
        // lane 4 f + b: exact bits of book b for family f (code lengths plus
        // category suffixes); per family the smallest total, ties to the
        // lowest book (native/src/codec/codec.c:1896-1922)
        uint64_t tot = 0u;
        if (lane < 12u) {
            const uint32_t f = lane >> 2;
            const uint32_t nsym = f == 0u ? 29u : (f == 1u ? 64u : 28u);
            const uint32_t base = f == 0u ? BH_DC : (f == 1u ? BH_RUN : BH_LVL);
            const uint32_t* len = K->vlc.len[lane];
            for (uint32_t s = 0; s < nsym; ++s) {
                const uint32_t c = sH[base + s];
                if (c != 0u) { tot += (uint64_t)c * (len[s] + (f != 1u && s > 1u ? s - 1u : 0u)); }
            }
        }
        uint32_t pick[3];
#pragma unroll
        for (uint32_t f = 0; f < 3u; ++f) {
            uint64_t best = readlane64(tot, 4u * f);
            uint32_t id = 0u;
            for (uint32_t b = 1; b < 4u; ++b) {
                const uint64_t t = readlane64(tot, 4u * f + b);
                if (t < best) { best = t; id = b; }
            }
            pick[f] = id;
        }
        k1 = pick[0];  // DC book
        k2 = pick[2];  // level book
        k3 = pick[1];  // run book
        sCode[0][lane] = K->vlc.code[k1][lane];
        sLen[0][lane] = K->vlc.len[k1][lane];
        sCode[1][lane] = K->vlc.code[4u + k3][lane];
        sLen[1][lane] = K->vlc.len[4u + k3][lane];
        sCode[2][lane] = K->vlc.code[8u + k2][lane];
        sLen[2][lane] = K->vlc.len[8u + k2][lane];
        __syncthreads();
    }
    k1 = wuni(k1); k2 = wuni(k2); k3 = wuni(k3);

    // ---------------- emit pass ----------------
    uint64_t veob;
    const uint32_t neob = vlc ? sLen[1][63] : rice_word(k3, 63u, veob);
    if (vlc) { veob = sCode[1][63]; }
    uint32_t* out = (uint32_t*)(scratch + S.scratch);
    const uint64_t cap_bits = (uint64_t)S.scratch_cap * 8u;
    uint64_t B = 0u;    // bits emitted
    uint32_t F = 0u;    // ring words stored
    uint32_t over = 0u;
    auto emit_block = [&](int32_t q, uint32_t dcm) {
        const uint64_t nz = __ballot(q != 0) & ~1ull;
        const uint32_t lm = min(tc_map_signed(q), 0x07FFFFFFu);
        const uint64_t bl = nz & below;
        const uint32_t pp = bl ? 63u - (uint32_t)__clzll(bl) : 0u;
        const uint32_t run = lane - pp - 1u;
        const bool active = lane != 0u && q != 0;
        uint32_t n1 = 0u, n2 = 0u;
        uint64_t v1 = 0u, v2 = 0u;
        if (lane == 0u) {
            n1 = vlc ? vlc_cat_word(sCode[0], sLen[0], dcm, v1) : rice_word(k1, dcm, v1);
        } else if (active) {
            if (vlc) {
                n1 = sLen[1][run];
                v1 = sCode[1][run];
                n2 = vlc_cat_word(sCode[2], sLen[2], lm, v2);
            } else {
                n1 = rice_word(k3, run, v1);
                n2 = rice_word(k2, lm, v2);
            }
        }
        // exclusive prefix sum of the lane bit counts
        const uint32_t bits = n1 + n2;
        uint32_t incl = bits;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
            if ((int)lane >= off) { incl += y; }
        }
        const uint32_t total = __builtin_amdgcn_readlane(incl, 63);
        if (B + total + neob > cap_bits) { over = 1u; return; }
        const uint64_t pos = B + (incl - bits);
        if (n1 != 0u) { ring_put(sRing, pos, n1, v1); }
        if (n2 != 0u) { ring_put(sRing, pos + n1, n2, v2); }
        if (lane == 0u) { ring_put(sRing, B + total, neob, veob); }
        B += total + neob;
        while ((uint32_t)(B >> 5) - F >= 64u) {
            const uint32_t w = F + lane;
            const uint32_t val = sRing[w & BW_MASK];
            sRing[w & BW_MASK] = 0u;
            out[w] = bswap32(val);
            F += 64u;
        }
    };
    {
        int32_t qa[BE_GROUP], qb[BE_GROUP];
        uint32_t da, db;
#pragma unroll
        for (uint32_t j = 0; j < BE_GROUP; ++j) { qa[j] = j < nblk ? qs[(size_t)j * 64u] : 0; }
        da = lane < BE_GROUP && lane < nblk ? dcmbuf[S.blk0 + lane] : 0u;
        for (uint32_t g = 0; g < nblk && !over; g += BE_GROUP) {
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) {
                const uint32_t b = g + BE_GROUP + j;
                qb[j] = b < nblk ? qs[(size_t)b * 64u] : 0;
            }
            db = lane < BE_GROUP && g + BE_GROUP + lane < nblk ? dcmbuf[S.blk0 + g + BE_GROUP + lane] : 0u;
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) {
                if (g + j < nblk && !over) { emit_block(qa[j], __builtin_amdgcn_readlane(da, j)); }
            }
#pragma unroll
            for (uint32_t j = 0; j < BE_GROUP; ++j) { qa[j] = qb[j]; }
            da = db;
        }
    }
    if (!over) {
        const uint32_t nw = (uint32_t)((B + 31u) >> 5);
        for (uint32_t w = F + lane; w < nw; w += 64u) { out[w] = bswap32(sRing[w & BW_MASK]); }
    }
    if (lane == 0u) {
        tcg_eresult R;
        R.start = 0u;
        R.size = over ? 0u : (uint32_t)((B + 7u) >> 3);
        R.status = any_viol ? -1 : (over ? -15 : 0);
        R.k1 = k1; R.k2 = k2; R.k3 = k3;
        R.crc = 0; R.offset = 0; R.aux = 0; R.pad_ = 0;
        results[blockIdx.x] = R;
    }
}

// ---------------------------------------------------------------------------
// K5b: V1 / V2 slice decoder, one wavefront per slice.
// The payload is read MSB-first through a 64-bit wave-uniform reservoir that
// is refilled with big-endian dwords from a 256-byte window held one dword
// per lane (the next window prefetched). Rice: the leading ones of the
// reservoir give the unary part in one count. VLC: lane l (1..20) holds the
// left-aligned end of the length-l code range; the lowest lane whose end
// exceeds the next 20 bits is the code length (ballot + find-first), and two
// lane reads give the symbol. The bits past the payload end are never
// trusted: a stream that runs over is rejected by the final length check.
// The payload must be readable 768 bytes past its end (window prefetch).
// ---------------------------------------------------------------------------
struct BitRd {
    uint64_t R;       // next payload bits, MSB-first, rc valid
    uint32_t rc;
    uint32_t widx;    // next dword of `cur` to move into the reservoir
    uint32_t nwin;    // windows fetched so far
    uint32_t maxwin;  // windows that may be fetched (payload end + 256)
    uint32_t loaded;  // payload bits moved into the reservoir
    uint32_t cur;     // lane i: dword i of the current window, big-endian
    uint32_t nxt;     // lane i: dword i of the next window, as stored
    gcu32* wp;        // address of the window after `nxt`
};

__device__ static inline void br_refill(BitRd& D, uint32_t lane)
{
    if (D.widx == 64u) {
        // past the last window that may hold payload bytes the same window is
        // read again; such bits only reach a stream the length check rejects
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

// at least 32 valid bits afterwards
__device__ static inline void br_need32(BitRd& D, uint32_t lane)
{
    if (D.rc < 32u) { br_refill(D, lane); }
}

__device__ static inline void br_init(BitRd& D, const uint8_t* payload, uint32_t n, uint32_t lane)
{
    const uintptr_t a = (uintptr_t)payload;
    const uintptr_t w0 = a & ~(uintptr_t)255u;
    gcu32* wp = (gcu32*)w0;
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

__device__ static inline void br_skip(BitRd& D, uint32_t n)
{
    D.R <<= n;
    D.rc -= n;
}

// top n bits (0 <= n <= 32) of the reservoir
__device__ static inline uint32_t br_peek(const BitRd& D, uint32_t n)
{
    return (uint32_t)((D.R >> 32) >> (32u - n));
}

// one Rice symbol (native/src/entropy/rice.h:160-246)
__device__ static inline uint32_t br_rice(BitRd& D, uint32_t k, uint32_t lane)
{
    br_need32(D, lane);
    // leading ones, 31 or more meaning the escape
    const uint32_t ones = (uint32_t)__builtin_clz(~(uint32_t)(D.R >> 32) | 1u);
    if (__builtin_expect(ones < 31u, 1)) {
        br_skip(D, ones + 1u);
        br_need32(D, lane);
        const uint32_t rem = br_peek(D, k);
        br_skip(D, k);
        return (ones << k) | rem;
    }
    br_skip(D, 31u);
    br_need32(D, lane);
    const uint32_t m = br_peek(D, 32u);
    br_skip(D, 32u);
    return m;
}

// one VLC code word; returns the symbol (native/src/entropy/vlc.c:108-126)
__device__ static inline uint32_t br_vlc(BitRd& D, uint32_t limv, uint32_t infov, uint32_t symv, uint32_t lane)
{
    br_need32(D, lane);
    const uint32_t top = br_peek(D, 20u);
    const uint64_t hit = __ballot(limv > top);
    // the frozen books are complete, so lane 20 (end 2^20) always matches
    const uint32_t L = hit ? (uint32_t)__builtin_ctzll(hit) : 20u;
    const uint32_t info = __builtin_amdgcn_readlane(infov, L);
    const uint32_t code = top >> (20u - L);
    const uint32_t i = ((info & 0xFFu) + code - (info >> 8)) & 63u;
    br_skip(D, L);
    return __builtin_amdgcn_readlane(symv, i);
}

// category symbol + (cat - 1) suffix bits (native/src/entropy/vlc.c:130-151)
__device__ static inline uint32_t br_vlc_cat(BitRd& D, uint32_t limv, uint32_t infov, uint32_t symv, uint32_t lane)
{
    const uint32_t cat = br_vlc(D, limv, infov, symv, lane);
    if (cat == 0u) { return 0u; }
    const uint32_t sb = cat - 1u;
    br_need32(D, lane);
    const uint32_t sfx = br_peek(D, sb);
    br_skip(D, sb);
    return (1u << sb) | sfx;
}

template <bool VLC>
__device__ static int32_t bits_decode_body(BitRd& D, const tcg_dslice& S, uint32_t cols, const uint32_t* tv,
                                           int32_t* __restrict__ qbuf, int32_t* sDcRow, uint32_t lane)
{
    // 12-bit symbol domains (native/src/entropy/rice.h:23-27,51-71)
    const uint32_t dc_max = 0x08000000u, lv_max = 0x04000000u;
    const int64_t dlim = (int64_t)1 << 25;
    const uint32_t myscan = tc_zigzag_pos(lane);
    // decoded blocks are stored column-major: natural index 8 r + c at 8 c + r
    int32_t* out = qbuf + (size_t)S.blk0 * 64u + ((lane & 7u) << 3 | lane >> 3);
    int32_t left = 0;
    for (uint32_t by = 0; by < S.h; ++by) {
        for (uint32_t bx = 0; bx < cols; ++bx) {
            const int32_t top_v = sDcRow[bx];
            const uint32_t dm = VLC ? br_vlc_cat(D, tv[0], tv[1], tv[2], lane) : br_rice(D, S.k1, lane);
            const int32_t top = by ? (int32_t)wuni((uint32_t)top_v) : 0;
            const int32_t pred = tc_dc_predict(bx != 0u, left, by != 0u, top);
            const int64_t dv = (int64_t)pred + (int64_t)tc_unmap_signed(dm);
            if (dm > dc_max || dv > dlim || dv < -dlim) { return -9; }
            const int32_t dc = (int32_t)dv;
            if (lane == 0u) { sDcRow[bx] = dc; }
            left = dc;
            uint32_t val = 0u;
            uint32_t pos = 1u;
            for (;;) {
                const uint32_t run = VLC ? br_vlc(D, tv[3], tv[4], tv[5], lane) : br_rice(D, S.k3, lane);
                if (run == 63u) { break; }
                const uint32_t idx = pos + run;
                if (run > 63u || idx > 63u) { return -9; }
                const uint32_t m = VLC ? br_vlc_cat(D, tv[6], tv[7], tv[8], lane) : br_rice(D, S.k2, lane);
                if (m > lv_max) { return -9; }
                val = myscan == idx ? m : val;
                pos = idx + 1u;
            }
            const int32_t v = lane == 0u ? dc : tc_unmap_signed(val);
            out[((size_t)by * cols + bx) * 64u] = v;
        }
    }
    return 0;
}

__global__ __launch_bounds__(64) void k_slice_decode_bits(const TcgConsts* __restrict__ K,
                                                           const tcg_dplane* __restrict__ planes,
                                                           const tcg_dslice* __restrict__ slices,
                                                           int32_t* __restrict__ qbuf,
                                                           int32_t* __restrict__ status)
{
    __shared__ uint32_t sTab[256];
    __shared__ uint32_t sX2n[32];
    extern __shared__ int32_t sDcRow[];  // one per block column of the batch's widest plane

    const uint32_t lane = threadIdx.x;
    const tcg_dslice S = slices[blockIdx.x];
    if (S.kind != 0u || S.mode == TCG_MODE_RANS2) { return; }
    const tcg_dplane P = planes[S.plane];
    for (uint32_t i = lane; i < 256u; i += 64u) { sTab[i] = K->crc[i]; }
    if (lane < 32u) { sX2n[lane] = K->x2n[lane]; }
    __syncthreads();

    const uint8_t* pl = (const uint8_t*)S.payload;
    const uint32_t n = S.size;
    int32_t rc = 0;
    if (wuni(wave_crc32(pl, n, sTab, sX2n)) != S.crc) { rc = -11; }
    const bool vlc = S.mode == TCG_MODE_VLC;
    if (rc == 0 && (vlc ? (S.k1 > 3u || S.k2 > 3u || S.k3 > 3u) : (S.k1 > 14u || S.k2 > 14u || S.k3 > 14u))) {
        rc = -9;
    }
    if (rc == 0) {
        BitRd D;
        br_init(D, pl, n, lane);
        if (vlc) {
            // lane tables of the three books: DC (k1), run (k3), level (k2)
            uint32_t tv[9];
            const uint32_t slot[3] = { S.k1, 4u + S.k3, 8u + S.k2 };
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                tv[3 * f + 0] = K->vlc.lim[slot[f]][lane];
                tv[3 * f + 1] = K->vlc.info[slot[f]][lane];
                tv[3 * f + 2] = K->vlc.sym[slot[f]][lane];
            }
            rc = bits_decode_body<true>(D, S, P.bcols, tv, qbuf, sDcRow, lane);
        } else {
            rc = bits_decode_body<false>(D, S, P.bcols, nullptr, qbuf, sDcRow, lane);
        }
        // byte alignment inside the payload, then exact consumption
        if (rc == 0 && ((D.loaded - D.rc + 7u) >> 3) != n) { rc = -9; }
    }
    if (lane == 0u) { status[blockIdx.x] = rc == 0 ? 0 : 1; }
}

// ---------------------------------------------------------------------------
// launches (queued beside the rANS2 kernels; each kernel skips the slices of
// the other modes)
// ---------------------------------------------------------------------------
void tcg_bits_encode_launch(hipStream_t st, const TcgConsts* K, const tcg_eplane* planes,
                            const tcg_eslice* slices, uint32_t nslices, const int32_t* qbuf, uint32_t* dcmbuf,
                            uint8_t* scratch, tcg_eresult* results, uint32_t max_cols)
{
    hipLaunchKernelGGL(k_slice_encode_bits, dim3(nslices), dim3(64), (size_t)max_cols * 4u, st, K, planes, slices,
                       qbuf, dcmbuf, scratch, results);
}

void tcg_bits_decode_launch(hipStream_t st, const TcgConsts* K, const tcg_dplane* planes,
                            const tcg_dslice* slices, uint32_t nslices, int32_t* qbuf, int32_t* status,
                            uint32_t max_cols)
{
    hipLaunchKernelGGL(k_slice_decode_bits, dim3(nslices), dim3(64), (size_t)max_cols * 4u, st, K, planes, slices,
                       qbuf, status);
}
