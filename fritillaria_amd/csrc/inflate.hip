// inflate.hip -- BGZF member inflate + CRC32 for gfx950, one wave64 per member.
//
// Behaviour follows crates/fritillaria-cuda/kernels/inflate.cu:23-458 (RFC 1951
// stored / fixed / dynamic blocks, its status taxonomy 0-9 and the order in
// which it detects errors, then CRC-32/ISO-HDLC of the produced bytes). The
// design is our own:
//
//  * The Huffman decode is wave-uniform: every lane holds the same bit-reader
//    state, so it lives in SGPRs and runs on the scalar unit, while the vector
//    lanes do the parallel parts -- table construction (ballot ranks, LUT fill
//    by a canonical walk per entry), LZ77 copies (one byte per lane per step),
//    the HBM write-back and the CRC.
//  * Output is staged in a 32 KiB LDS ring (DEFLATE's window is 32 KiB), so
//    back-references never touch HBM; it is flushed to HBM every ~8 KB.
//  * CRC is folded in during each flush: lane l takes a 132-byte segment
//    (stride 33 dwords: conflict-free), and the segments are combined with a
//    per-lane GF(2) factor x^(8*132*(63-l)) and a wave XOR reduction.
//  * Compressed input is staged through a 2 KiB LDS window fed by coalesced
//    16-byte loads, so the scalar refill never waits on HBM per symbol.
//  * Persistent form: a grid of resident waves pulls member indices from a
//    device work-queue counter (optionally through a size-sorted order).
//
// LDS per wave: 39.8 KB -> 4 waves per CU.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crc_math.h"
#include "frhip_internal.h"
#include "host_internal.h"

namespace frhip {
namespace {

constexpr uint32_t RING = 32768, RMASK = RING - 1;
constexpr uint32_t STAGE_BYTES = 2048, STAGE_WORDS = STAGE_BYTES / 4, CHUNK = 1024;
constexpr uint32_t FLUSH_AT = 8000;
constexpr uint32_t SEG = 132; // 64 * 132 = 8448 >= FLUSH_AT + 258
constexpr int LBITS = 10, DBITS = 8, CBITS = 7;

enum : uint32_t {
    ST_OK = 0,
    ST_TRUNCATED = 1,
    ST_BLOCK_TYPE = 2,
    ST_STORED_LEN = 3,
    ST_OUTPUT_OVERFLOW = 4,
    ST_BAD_CODE = 5,
    ST_BAD_DISTANCE = 6,
    ST_BAD_TABLE = 7,
    ST_BAD_LENGTH_REPEAT = 8,
    ST_BAD_COUNTS = 9,
};

__constant__ uint16_t LEN_BASE[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t DIST_BASE[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,    25,
                                       33,   49,   65,   97,   129,  193,   257,   385,   513,   769,
                                       1025, 1537, 2049, 3073, 4097, 6145,  8193,  12289, 16385, 24577};
__constant__ uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
__constant__ uint8_t CLEN_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

struct Smem {
    uint8_t ring[RING];
    uint32_t stage[STAGE_WORDS];
    uint16_t llut[1 << LBITS];
    uint16_t dlut[1 << DBITS];
    uint16_t clut[1 << CBITS];
    uint16_t lsym[288];
    uint16_t dsym[32];
    uint16_t csym[20];
    uint16_t lcount[16];
    uint16_t dcount[16];
    uint16_t ccount[16];
    uint8_t lens[320];
    uint32_t crctab[256];
};

#define WFENCE() __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront")

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Wave-uniform LSB-first bit reader over one member's payload.
struct Bits {
    uint64_t buf;
    uint32_t cnt;       // valid bits in buf (never counts bytes past the payload)
    uint32_t pos;       // payload bytes moved into buf
    uint32_t len;       // payload length
    uint64_t base;      // absolute offset of the payload in the batch buffer
    uint64_t stage_end; // absolute end of the staged window (multiple of CHUNK)
};

__device__ __forceinline__ void load_chunk(Smem &sm, const uint8_t *gin, uint64_t chunk_abs, uint32_t lane)
{
    const uint4 v = *reinterpret_cast<const uint4 *>(gin + chunk_abs + 16u * lane);
    reinterpret_cast<uint4 *>(sm.stage)[((chunk_abs & (STAGE_BYTES - 1)) >> 4) + lane] = v;
}

// Makes bytes [abs & ~3, (abs & ~3) + 12) resident in the stage.
__device__ __forceinline__ void ensure_stage(Smem &sm, Bits &b, const uint8_t *gin, uint64_t abs, uint32_t lane)
{
    uint64_t need_lo = abs & ~3ull, need_hi = need_lo + 12;
    if (need_hi <= b.stage_end && need_lo + STAGE_BYTES >= b.stage_end) return;
    if (need_lo + CHUNK >= b.stage_end && need_hi <= b.stage_end + CHUNK && b.stage_end != 0) {
        load_chunk(sm, gin, b.stage_end, lane);
        b.stage_end += CHUNK;
    } else {
        uint64_t c0 = abs & ~(uint64_t)(CHUNK - 1);
        load_chunk(sm, gin, c0, lane);
        load_chunk(sm, gin, c0 + CHUNK, lane);
        b.stage_end = c0 + 2 * CHUNK;
    }
    WFENCE();
}

__device__ __forceinline__ void refill(Smem &sm, Bits &b, const uint8_t *gin, uint32_t lane)
{
    if (b.cnt > 56) return;
    uint32_t avail = b.len - b.pos;
    if (avail == 0) return;
    uint32_t nbytes = (64u - b.cnt) >> 3;
    if (nbytes > avail) nbytes = avail;
    uint64_t abs = b.base + b.pos;
    ensure_stage(sm, b, gin, abs, lane);
    uint32_t w = (uint32_t)(abs >> 2) & (STAGE_WORDS - 1);
    uint32_t w0 = uni(sm.stage[w]);
    uint32_t w1 = uni(sm.stage[(w + 1) & (STAGE_WORDS - 1)]);
    uint32_t w2 = uni(sm.stage[(w + 2) & (STAGE_WORDS - 1)]);
    uint64_t lo = ((uint64_t)w1 << 32) | w0;
    uint32_t sh = (uint32_t)(abs & 3) * 8;
    uint64_t v = sh ? ((lo >> sh) | ((uint64_t)w2 << (64 - sh))) : lo;
    if (nbytes < 8) v &= (1ull << (8 * nbytes)) - 1;
    b.buf |= v << b.cnt;
    b.cnt += 8 * nbytes;
    b.pos += nbytes;
}

__device__ __forceinline__ uint32_t getbits(Bits &b, uint32_t n, uint32_t &err)
{
    if (n > b.cnt) {
        err = ST_TRUNCATED;
        return 0;
    }
    uint32_t v = (uint32_t)(b.buf & ((1ull << n) - 1));
    b.buf >>= n;
    b.cnt -= n;
    return v;
}

// One symbol through a LUT of `lbits`, falling back to the canonical walk for
// codes longer than the LUT and for invalid codes (puff semantics: an invalid
// code is reported after 15 bits, a stream that ends first is truncated).
template <int LB>
__device__ __forceinline__ int decode_sym(Bits &b, const uint16_t *lut, const uint16_t *count,
                                          const uint16_t *sym, uint32_t bad, uint32_t &err)
{
    uint32_t e = uni(lut[b.buf & ((1u << LB) - 1)]);
    if (e) {
        uint32_t L = e & 15;
        if (L > b.cnt) {
            err = ST_TRUNCATED;
            return -1;
        }
        b.buf >>= L;
        b.cnt -= L;
        return (int)(e >> 4);
    }
    int code = 0, first = 0, index = 0;
    for (int len = 1; len <= 15; ++len) {
        if ((uint32_t)len > b.cnt) {
            err = ST_TRUNCATED;
            return -1;
        }
        code |= (int)((b.buf >> (len - 1)) & 1);
        int c = (int)uni(count[len]);
        if (code - first < c) {
            b.buf >>= len;
            b.cnt -= len;
            return (int)uni(sym[index + code - first]);
        }
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    err = bad;
    return -1;
}

// Canonical code construction from lengths (puff's construct, lane-parallel):
// counts by ballot, symbol order by ballot ranks. Returns 0 complete, >0
// incomplete, <0 over-subscribed; *count0 = symbols of length 0.
__device__ int construct(const uint8_t *lens, int n, uint16_t *count_out, uint16_t *sym_out, uint32_t lane,
                         int *count0)
{
    int cnt[16];
#pragma unroll
    for (int l = 0; l < 16; ++l) cnt[l] = 0;
    for (int r = 0; r < n; r += 64) {
        int s = r + (int)lane;
        int L = s < n ? (int)lens[s] : -1;
#pragma unroll
        for (int l = 0; l < 16; ++l) cnt[l] += __popcll(__ballot(L == l));
    }
    if (lane == 0) {
#pragma unroll
        for (int l = 0; l < 16; ++l) count_out[l] = (uint16_t)cnt[l];
    }
    *count0 = cnt[0];
    if (cnt[0] == n) {
        WFENCE();
        return 0;
    }
    int left = 1;
#pragma unroll
    for (int l = 1; l < 16; ++l) {
        left <<= 1;
        left -= cnt[l];
        if (left < 0) {
            WFENCE();
            return left;
        }
    }
    int offs[16];
    offs[1] = 0;
#pragma unroll
    for (int l = 1; l < 15; ++l) offs[l + 1] = offs[l] + cnt[l];
    const uint64_t lt = (1ull << lane) - 1;
    for (int r = 0; r < n; r += 64) {
        int s = r + (int)lane;
        int L = s < n ? (int)lens[s] : 0;
#pragma unroll
        for (int l = 1; l < 16; ++l) {
            uint64_t m = __ballot(L == l);
            if (L == l) sym_out[offs[l] + __popcll(m & lt)] = (uint16_t)s;
            offs[l] += __popcll(m);
        }
    }
    WFENCE();
    return left;
}

// LUT entry = (symbol << 4) | length for codes of at most LB bits, else 0.
template <int LB>
__device__ void build_lut(uint16_t *lut, const uint16_t *count, const uint16_t *sym, uint32_t lane)
{
    int cnt[LB + 1];
#pragma unroll
    for (int l = 1; l <= LB; ++l) cnt[l] = (int)uni(count[l]);
    for (int e = (int)lane; e < (1 << LB); e += 64) {
        int code = 0, first = 0, index = 0;
        uint16_t out = 0;
        bool found = false;
#pragma unroll
        for (int len = 1; len <= LB; ++len) {
            if (!found) {
                code |= (e >> (len - 1)) & 1;
                if (code - first < cnt[len]) {
                    out = (uint16_t)((sym[index + code - first] << 4) | len);
                    found = true;
                } else {
                    index += cnt[len];
                    first = (first + cnt[len]) << 1;
                    code <<= 1;
                }
            }
        }
        lut[e] = out;
    }
    WFENCE();
}

struct Out {
    uint8_t *g;         // member's output in HBM
    uint32_t pos;       // bytes produced
    uint32_t flushed;   // bytes written back to HBM (and folded into the CRC)
    uint32_t cap;       // ISIZE
    uint32_t raw;       // zero-init CRC register of bytes [0, flushed)
};

// Writes ring bytes [flushed, pos) to HBM and folds them into the CRC.
__device__ void flush(Smem &sm, Out &o, uint32_t lane, uint32_t lane_factor, const uint32_t *x2n)
{
    uint32_t c = o.pos - o.flushed;
    if (c == 0) return;
    WFENCE();
    for (uint32_t i = lane; i < c; i += 64) o.g[o.flushed + i] = sm.ring[(o.flushed + i) & RMASK];
    // segment of lane l: [c - SEG*(64-l), c - SEG*(63-l)) clipped at 0
    int32_t hi = (int32_t)c - (int32_t)(SEG * (63 - lane));
    int32_t lo = hi - (int32_t)SEG;
    if (lo < 0) lo = 0;
    uint32_t r = 0;
    for (int32_t k = lo; k < hi; ++k) {
        uint32_t byte = sm.ring[(o.flushed + (uint32_t)k) & RMASK];
        r = sm.crctab[(r ^ byte) & 0xFF] ^ (r >> 8);
    }
    r = fr_mulmod(r, lane_factor);
    // XOR reduction across the wave
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r ^= __shfl_xor(r, off, 64);
    uint32_t chunk = uni(r);
    o.raw = fr_mulmod(o.raw, fr_xpow8(x2n, c)) ^ chunk;
    o.flushed = o.pos;
}

__device__ void inflate_member(Smem &sm, const uint8_t *gin, uint64_t in_off, uint32_t in_len, uint8_t *gout,
                               uint32_t cap, uint32_t lane, uint32_t lane_factor, const uint32_t *x2n,
                               uint32_t &status_out, uint32_t &produced_out, uint32_t &crc_out)
{
    Bits b;
    b.buf = 0;
    b.cnt = 0;
    b.pos = 0;
    b.len = in_len;
    b.base = in_off;
    b.stage_end = 0;
    Out o;
    o.g = gout;
    o.pos = 0;
    o.flushed = 0;
    o.cap = cap;
    o.raw = 0;
    uint32_t err = ST_OK;
    int tables = 0; // 0 none, 1 fixed, 2 dynamic
    uint32_t last;
    do {
        refill(sm, b, gin, lane);
        last = getbits(b, 1, err);
        uint32_t type = getbits(b, 2, err);
        if (err) break;
        if (type == 0) {
            // stored: drop to a byte boundary, give back whole buffered bytes
            b.pos -= b.cnt >> 3;
            b.buf = 0;
            b.cnt = 0;
            if (b.pos + 4 > b.len) {
                err = ST_TRUNCATED;
                break;
            }
            const uint8_t *p = gin + b.base + b.pos;
            uint32_t len = uni((uint32_t)p[0] | ((uint32_t)p[1] << 8));
            uint32_t nlen = uni((uint32_t)p[2] | ((uint32_t)p[3] << 8));
            b.pos += 4;
            if (len != (~nlen & 0xffffu)) {
                err = ST_STORED_LEN;
                break;
            }
            if (b.pos + len > b.len) {
                err = ST_TRUNCATED;
                break;
            }
            if (o.pos + len > o.cap) {
                err = ST_OUTPUT_OVERFLOW;
                break;
            }
            const uint8_t *src = gin + b.base + b.pos;
            for (uint32_t done = 0; done < len; done += 256) {
                if (o.pos - o.flushed >= FLUSH_AT) flush(sm, o, lane, lane_factor, x2n);
                uint32_t piece = len - done < 256 ? len - done : 256;
                for (uint32_t i = lane; i < piece; i += 64) sm.ring[(o.pos + i) & RMASK] = src[done + i];
                o.pos += piece;
            }
            b.pos += len;
            continue;
        }
        if (type == 3) {
            err = ST_BLOCK_TYPE;
            break;
        }
        if (type == 1) {
            if (tables != 1) {
                for (uint32_t s = lane; s < 288; s += 64)
                    sm.lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : 8;
                WFENCE();
                int c0;
                construct(sm.lens, 288, sm.lcount, sm.lsym, lane, &c0);
                if (lane < 32) sm.lens[lane] = 5;
                WFENCE();
                construct(sm.lens, 32, sm.dcount, sm.dsym, lane, &c0);
                build_lut<LBITS>(sm.llut, sm.lcount, sm.lsym, lane);
                build_lut<DBITS>(sm.dlut, sm.dcount, sm.dsym, lane);
                tables = 1;
            }
        } else {
            // dynamic header
            uint32_t nlen = getbits(b, 5, err) + 257;
            uint32_t ndist = getbits(b, 5, err) + 1;
            uint32_t ncode = getbits(b, 4, err) + 4;
            if (err) break;
            if (nlen > 286 || ndist > 30) {
                err = ST_BAD_COUNTS;
                break;
            }
            if (lane < 19) sm.lens[lane] = 0;
            WFENCE();
            for (uint32_t i = 0; i < ncode; ++i) {
                if (b.cnt < 3) refill(sm, b, gin, lane);
                uint32_t v = getbits(b, 3, err);
                if (err) break;
                if (lane == 0) sm.lens[CLEN_ORDER[i]] = (uint8_t)v;
            }
            if (err) break;
            WFENCE();
            int c0;
            if (construct(sm.lens, 19, sm.ccount, sm.csym, lane, &c0) != 0) {
                err = ST_BAD_TABLE;
                break;
            }
            build_lut<CBITS>(sm.clut, sm.ccount, sm.csym, lane);
            uint32_t index = 0, total = nlen + ndist;
            while (index < total) {
                refill(sm, b, gin, lane);
                int sym = decode_sym<CBITS>(b, sm.clut, sm.ccount, sm.csym, ST_BAD_CODE, err);
                if (err) break;
                if (sym < 16) {
                    if (lane == 0) sm.lens[index] = (uint8_t)sym;
                    ++index;
                    continue;
                }
                uint32_t len = 0, rep;
                if (sym == 16) {
                    if (index == 0) {
                        err = ST_BAD_LENGTH_REPEAT;
                        break;
                    }
                    WFENCE();
                    len = uni(sm.lens[index - 1]);
                    rep = 3 + getbits(b, 2, err);
                } else if (sym == 17) {
                    rep = 3 + getbits(b, 3, err);
                } else {
                    rep = 11 + getbits(b, 7, err);
                }
                if (err) break;
                if (index + rep > total) {
                    err = ST_BAD_LENGTH_REPEAT;
                    break;
                }
                for (uint32_t i = lane; i < rep; i += 64) sm.lens[index + i] = (uint8_t)len;
                index += rep;
            }
            if (err) break;
            WFENCE();
            if (uni(sm.lens[256]) == 0) {
                err = ST_BAD_TABLE;
                break;
            }
            int e = construct(sm.lens, (int)nlen, sm.lcount, sm.lsym, lane, &c0);
            if (e != 0 && (e < 0 || (int)nlen != c0 + 1)) {
                err = ST_BAD_TABLE;
                break;
            }
            e = construct(sm.lens + nlen, (int)ndist, sm.dcount, sm.dsym, lane, &c0);
            if (e != 0 && (e < 0 || (int)ndist != c0 + 1)) {
                err = ST_BAD_TABLE;
                break;
            }
            build_lut<LBITS>(sm.llut, sm.lcount, sm.lsym, lane);
            build_lut<DBITS>(sm.dlut, sm.dcount, sm.dsym, lane);
            tables = 2;
        }
        // literal / length / distance codes until end-of-block
        for (;;) {
            if (o.pos - o.flushed >= FLUSH_AT) flush(sm, o, lane, lane_factor, x2n);
            refill(sm, b, gin, lane);
            int sym = decode_sym<LBITS>(b, sm.llut, sm.lcount, sm.lsym, ST_BAD_CODE, err);
            if (err) break;
            if (sym < 256) {
                if (o.pos >= o.cap) {
                    err = ST_OUTPUT_OVERFLOW;
                    break;
                }
                if (lane == 0) sm.ring[o.pos & RMASK] = (uint8_t)sym;
                ++o.pos;
                continue;
            }
            if (sym == 256) break;
            sym -= 257;
            if (sym >= 29) {
                err = ST_BAD_CODE;
                break;
            }
            uint32_t len = LEN_BASE[sym] + getbits(b, LEN_EXTRA[sym], err);
            if (err) break;
            int ds = decode_sym<DBITS>(b, sm.dlut, sm.dcount, sm.dsym, ST_BAD_DISTANCE, err);
            if (err) break;
            if (ds >= 30) {
                err = ST_BAD_DISTANCE;
                break;
            }
            uint32_t dist = DIST_BASE[ds] + getbits(b, DIST_EXTRA[ds], err);
            if (err) break;
            if (dist > o.pos) {
                err = ST_BAD_DISTANCE;
                break;
            }
            if (o.pos + len > o.cap) {
                err = ST_OUTPUT_OVERFLOW;
                break;
            }
            WFENCE();
            uint32_t src0 = o.pos - dist;
            if (dist >= len) {
                for (uint32_t off = lane; off < len; off += 64)
                    sm.ring[(o.pos + off) & RMASK] = sm.ring[(src0 + off) & RMASK];
            } else {
                // period-`dist` pattern: every source byte precedes o.pos
                uint32_t r = lane % dist;
                uint32_t step = 64u % dist;
                for (uint32_t off = lane; off < len; off += 64) {
                    sm.ring[(o.pos + off) & RMASK] = sm.ring[(src0 + r) & RMASK];
                    r += step;
                    if (r >= dist) r -= dist;
                }
            }
            WFENCE();
            o.pos += len;
        }
        if (err) break;
    } while (!last);

    if (err == ST_OK) {
        flush(sm, o, lane, lane_factor, x2n);
        crc_out = o.raw ^ fr_mulmod(0xFFFFFFFFu, fr_xpow8(x2n, o.pos)) ^ 0xFFFFFFFFu;
    } else {
        crc_out = 0;
    }
    status_out = err;
    produced_out = o.pos;
}

__global__ __launch_bounds__(64) void fr_inflate_kernel(InflateArgs a)
{
    __shared__ Smem sm;
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 256; i += 64) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        sm.crctab[i] = c;
    }
    const uint32_t lane_factor = fr_xpow8(a.x2n, (uint64_t)SEG * (63 - lane));
    WFENCE();
    uint32_t iter = 0;
    for (;;) {
        uint32_t m;
        if (a.queue) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(a.queue, 1u);
            m = uni(t);
        } else {
            m = blockIdx.x + iter * gridDim.x;
        }
        ++iter;
        if (m >= a.n) break;
        if (a.order) m = a.order[m];
        uint32_t st, produced, crc;
        inflate_member(sm, a.in, a.in_off[m], a.in_len[m], a.out + a.out_off[m], a.out_cap[m], lane,
                       lane_factor, a.x2n, st, produced, crc);
        if (lane == 0) {
            a.status[m] = st;
            a.produced[m] = produced;
            a.crc[m] = crc;
        }
    }
}

// CRC32 of independent byte ranges, one wave per range: lane l takes the l-th
// of 64 equal segments; segments fold with x^(8 * bytes after the segment).
__global__ __launch_bounds__(64) void fr_crc32_kernel(const uint8_t *data, const uint64_t *off,
                                                      const uint32_t *len, uint32_t n, uint32_t *crc,
                                                      const uint32_t *x2n)
{
    __shared__ uint32_t tab[256];
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < 256; i += 64) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        tab[i] = c;
    }
    __syncthreads();
    for (uint32_t m = blockIdx.x; m < n; m += gridDim.x) {
        const uint8_t *p = data + off[m];
        const uint32_t L = len[m];
        const uint32_t S = (L + 63) / 64;
        uint32_t lo = lane * S, hi = lo + S;
        if (lo > L) lo = L;
        if (hi > L) hi = L;
        uint32_t r = 0;
        for (uint32_t k = lo; k < hi; ++k) r = tab[(r ^ p[k]) & 0xFF] ^ (r >> 8);
        r = fr_mulmod(r, fr_xpow8(x2n, L - hi));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) r ^= __shfl_xor(r, o, 64);
        if (lane == 0) crc[m] = r ^ fr_mulmod(0xFFFFFFFFu, fr_xpow8(x2n, L)) ^ 0xFFFFFFFFu;
    }
}

} // namespace

hipError_t launch_crc32_blocks(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n,
                               uint32_t *crc, const uint32_t *x2n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint32_t grid = n < 65536 ? n : 65536;
    hipLaunchKernelGGL(fr_crc32_kernel, dim3(grid), dim3(64), 0, s, data, off, len, n, crc, x2n);
    return hipGetLastError();
}

hipError_t launch_inflate(const InflateArgs &a, uint32_t grid, hipStream_t stream)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(fr_inflate_kernel, dim3(grid), dim3(64), 0, stream, a);
    return hipGetLastError();
}

} // namespace frhip
