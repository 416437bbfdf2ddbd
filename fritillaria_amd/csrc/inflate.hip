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
//    the HBM write-back and the CRC. Length and distance bases are computed,
//    not looked up, so the symbol loop touches only LDS.
//  * Output is staged in an LDS ring holding the last RING bytes of the member
//    (RING = 4..32 KiB by schedule). Ring slots follow the absolute output
//    offset, so write-back moves aligned 16-byte words to HBM. A back-reference
//    that reaches past the ring is read back from HBM, where this wave already
//    wrote it (after a workgroup-scope fence).
//  * CRC is folded in during each write-back: lane l takes a SEG-byte segment
//    and the segments are combined with a per-lane GF(2) factor
//    x^(8*SEG*(63-l)) and a wave XOR reduction.
//  * Compressed input is staged through a 1 KiB LDS window fed by coalesced
//    8-byte loads, so the scalar refill never waits on HBM per symbol.
//  * Persistent form: resident waves pull member indices from a device
//    work-queue counter (optionally through a size-sorted order).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "crc_math.h"
#include "frhip_internal.h"
#include "host_internal.h"

namespace frhip {
namespace {

constexpr uint32_t STAGE_BYTES = 1024, STAGE_WORDS = STAGE_BYTES / 4, CHUNK = 512;
constexpr int PBITS = 9, DBITS = 8, CBITS = 7;
// Literal/length table entry (PBITS-bit index): bits 0-3 length of the first
// code (0: longer than PBITS, take the canonical walk), 4-12 its symbol; when
// the first symbol is a literal and a second literal's code also fits in the
// index, bit 13 is set, bits 14-17 hold both codes' length and 18-25 the
// second literal.
constexpr uint32_t PAIR = 1u << 13;

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

// RFC 1951 3.2.5 in closed form. Length code c = sym - 257 (0..28): codes 0-7
// carry no extra bits (lengths 3-10), then each group of four doubles the step;
// code 28 is length 258. Distance code d (0..29): codes 0-3 are distances 1-4,
// then each pair doubles the step.
// Written as arithmetic and selects of computed values (no branches): codes
// 0-7 and 0-3 take the general form with no extra bits and a fixed offset.
__device__ __forceinline__ uint32_t len_extra(uint32_t c)
{
    const uint32_t x = (max(c, 4u) - 4) >> 2;
    return c == 28 ? 0 : x;
}
__device__ __forceinline__ uint32_t len_base(uint32_t c)
{
    const uint32_t g = ((4u | (c & 3)) << ((max(c, 4u) - 4) >> 2)) + 3 - (c < 4 ? 4u : 0u);
    return c == 28 ? 258 : g;
}
__device__ __forceinline__ uint32_t dist_extra(uint32_t d) { return (max(d, 2u) - 2) >> 1; }
__device__ __forceinline__ uint32_t dist_base(uint32_t d)
{
    return ((2u | (d & 1)) << dist_extra(d)) + 1 - (d < 2 ? 2u : 0u);
}
// Order of the code-length code lengths (RFC 1951 3.2.7), 5 bits per entry.
__device__ __forceinline__ uint32_t clen_order(uint32_t i)
{
    const uint64_t lo = (16ull << 0) | (17ull << 5) | (18ull << 10) | (0ull << 15) | (8ull << 20) | (7ull << 25) |
                        (9ull << 30) | (6ull << 35) | (10ull << 40) | (5ull << 45) | (11ull << 50) | (4ull << 55);
    const uint64_t hi = (12ull << 0) | (3ull << 5) | (13ull << 10) | (2ull << 15) | (14ull << 20) | (1ull << 25) |
                        (15ull << 30);
    return i < 12 ? (uint32_t)(lo >> (5 * i)) & 31 : (uint32_t)(hi >> (5 * (i - 12))) & 31;
}

// Per-wave LDS: the output ring, the staged compressed input and the Huffman
// tables of the current block.
template <uint32_t RING>
struct alignas(16) WaveMem {
    uint8_t ring[RING];
    uint32_t stage[STAGE_WORDS];
    uint32_t plut[1 << PBITS];
    uint16_t dlut[1 << DBITS];
    uint16_t clut[1 << CBITS];
    uint16_t lsym[288];
    uint16_t dsym[32];
    uint16_t csym[20];
    uint16_t lcount[16];
    uint16_t dcount[16];
    uint16_t ccount[16];
    uint32_t tok[64];   // decoded, not yet resolved symbols (token form)
    uint8_t lens[320];
};

#define WFENCE() __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront")

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    // the builtin returns int: each half is narrowed to uint32_t before it is
    // widened, or the low half would be sign-extended over the high one
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | lo;
}

// Wave-uniform LSB-first bit reader over one member's payload. Offsets are
// relative to `sbase`, the payload start rounded down to CHUNK.
struct Bits {
    uint64_t buf;
    uint32_t cnt;       // valid bits in buf (never counts bytes past the payload)
    uint32_t pos;       // payload bytes moved into buf (relative to the payload start)
    uint32_t len;       // payload length
    uint32_t lead;      // payload start - sbase
    uint32_t stage_end; // end of the staged window, relative to sbase (multiple of CHUNK)
    const uint8_t *sbase;
};

__device__ __forceinline__ void load_chunk(uint32_t *stage, const uint8_t *sbase, uint32_t rel, uint32_t lane)
{
    const uint2 v = *reinterpret_cast<const uint2 *>(sbase + rel + 8u * lane);
    reinterpret_cast<uint2 *>(stage)[((rel & (STAGE_BYTES - 1)) >> 3) + lane] = v;
}

// Makes bytes [r & ~7, (r & ~7) + 16) resident in the stage (r relative to sbase).
__device__ __forceinline__ void ensure_stage(uint32_t *stage, Bits &b, uint32_t r, uint32_t lane)
{
    uint32_t lo = r & ~7u, hi = lo + 16;
    if (hi <= b.stage_end && lo + STAGE_BYTES >= b.stage_end) return;
    if (b.stage_end != 0 && lo + CHUNK >= b.stage_end && hi <= b.stage_end + CHUNK) {
        load_chunk(stage, b.sbase, b.stage_end, lane);
        b.stage_end += CHUNK;
    } else {
        uint32_t c0 = r & ~(CHUNK - 1);
        load_chunk(stage, b.sbase, c0, lane);
        load_chunk(stage, b.sbase, c0 + CHUNK, lane);
        b.stage_end = c0 + 2 * CHUNK;
    }
    WFENCE();
}

// Tops the bit buffer up to at least 56 bits (while payload remains).
__device__ __forceinline__ void refill(uint32_t *stage, Bits &b, uint32_t lane)
{
    uint32_t avail = b.len - b.pos;
    if (avail == 0) return;
    uint32_t nbytes = (64u - b.cnt) >> 3;
    if (nbytes > avail) nbytes = avail;
    uint32_t r = b.lead + b.pos;
    ensure_stage(stage, b, r, lane);
    // two aligned 8-byte reads cover [r, r + 8)
    const uint32_t q = (r >> 3) & (STAGE_WORDS / 2 - 1);
    const uint2 x = reinterpret_cast<const uint2 *>(stage)[q];
    const uint2 y = reinterpret_cast<const uint2 *>(stage)[(q + 1) & (STAGE_WORDS / 2 - 1)];
    const uint64_t lo = ((uint64_t)uni(x.y) << 32) | uni(x.x);
    const uint64_t hi = ((uint64_t)uni(y.y) << 32) | uni(y.x);
    uint32_t sh = (r & 7) * 8;
    uint64_t v = sh ? ((lo >> sh) | (hi << (64 - sh))) : lo;
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

// One symbol through a LUT of LB bits, falling back to the canonical walk for
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

// Canonical code construction from lengths (puff's construct), lane-parallel.
// Lane l keeps the count of length l in a VGPR; symbols are placed by ballot
// ranks, one distinct length at a time. Returns 0 complete, >0 incomplete, <0
// over-subscribed; *count0 = symbols of length 0.
__device__ int construct(const uint8_t *lens, int n, uint16_t *count_out, uint16_t *sym_out, uint32_t lane,
                         int *count0)
{
    uint32_t cntv = 0; // lane l: symbols of length l
    for (int r = 0; r < n; r += 64) {
        int s = r + (int)lane;
        int L = s < n ? (int)lens[s] : 16;
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            uint32_t c = (uint32_t)__popcll(__ballot(L == l));
            if ((int)lane == l) cntv += c;
        }
    }
    if (lane < 16) count_out[lane] = (uint16_t)cntv;
    const int c0 = (int)__builtin_amdgcn_readlane(cntv, 0);
    *count0 = c0;
    if (c0 == n) {
        WFENCE();
        return 0;
    }
    int left = 1;
#pragma unroll
    for (int l = 1; l < 16; ++l) {
        left <<= 1;
        left -= (int)__builtin_amdgcn_readlane(cntv, l);
        if (left < 0) {
            WFENCE();
            return left;
        }
    }
    // lane l: first index of length l in the sorted symbol table
    uint32_t offv;
    {
        uint32_t x = lane >= 1 && lane < 16 ? cntv : 0;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            uint32_t y = __shfl_up(x, d, 64);
            if ((int)lane >= d) x += y;
        }
        offv = __shfl_up(x, 1, 64);
        if (lane <= 1) offv = 0;
    }
    const uint64_t lt = (1ull << lane) - 1;
    for (int r = 0; r < n; r += 64) {
        int s = r + (int)lane;
        int L = s < n ? (int)lens[s] : 0;
        uint64_t todo = __ballot(L != 0);
        while (todo) {
            uint32_t lead = (uint32_t)__ffsll((unsigned long long)todo) - 1;
            uint32_t Ll = __builtin_amdgcn_readlane((uint32_t)L, lead);
            uint64_t m = __ballot(L == (int)Ll);
            uint32_t base = __builtin_amdgcn_readlane(offv, Ll);
            if (L == (int)Ll) sym_out[base + __popcll(m & lt)] = (uint16_t)s;
            if (lane == Ll) offv += (uint32_t)__popcll(m);
            todo &= ~m;
        }
    }
    WFENCE();
    return left;
}

// LUT entry = (symbol << 4) | length for codes of at most LB bits, else 0.
template <int LB>
__device__ void build_lut(uint16_t *lut, const uint16_t *count, const uint16_t *sym, uint32_t lane)
{
    uint32_t cv = lane < 16 ? count[lane] : 0;
    for (int e = (int)lane; e < (1 << LB); e += 64) {
        int code = 0, first = 0, index = 0;
        uint16_t out = 0;
        bool found = false;
#pragma unroll
        for (int len = 1; len <= LB; ++len) {
            int c = (int)__builtin_amdgcn_readlane(cv, len);
            if (!found) {
                code |= (e >> (len - 1)) & 1;
                if (code - first < c) {
                    out = (uint16_t)((sym[index + code - first] << 4) | len);
                    found = true;
                } else {
                    index += c;
                    first = (first + c) << 1;
                    code <<= 1;
                }
            }
        }
        lut[e] = out;
    }
    WFENCE();
}

// First code of `e`'s entry (see PAIR), or the canonical walk for codes longer
// than PBITS. Same results and error order as decode_sym.
__device__ __forceinline__ int decode_litlen(Bits &b, uint32_t e, const uint16_t *count, const uint16_t *sym,
                                             uint32_t &err)
{
    if (e & 15) {
        const uint32_t L = e & 15;
        if (L > b.cnt) {
            err = ST_TRUNCATED;
            return -1;
        }
        b.buf >>= L;
        b.cnt -= L;
        return (int)((e >> 4) & 511);
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
    err = ST_BAD_CODE;
    return -1;
}

// Canonical walk of the low bits of `bits`, at most `maxlen` of them; lane l of
// cv holds the count of codes of length l. Returns length (0: none) and symbol.
__device__ __forceinline__ uint32_t walk_code(uint32_t bits, uint32_t maxlen, uint32_t cv, const uint16_t *sym,
                                              uint32_t &out_sym)
{
    int code = 0, first = 0, index = 0;
    uint32_t found = 0;
#pragma unroll
    for (int len = 1; len <= PBITS; ++len) {
        int c = (int)__builtin_amdgcn_readlane(cv, len);
        if (!found && (uint32_t)len <= maxlen) {
            code |= (int)((bits >> (len - 1)) & 1);
            if (code - first < c) {
                out_sym = sym[index + code - first];
                found = (uint32_t)len;
            } else {
                index += c;
                first = (first + c) << 1;
                code <<= 1;
            }
        }
    }
    return found;
}

// The literal/length table with literal pairs (see PAIR).
__device__ void build_pair_lut(uint32_t *lut, const uint16_t *count, const uint16_t *sym, uint32_t lane)
{
    uint32_t cv = lane < 16 ? count[lane] : 0;
    for (uint32_t e = lane; e < (1u << PBITS); e += 64) {
        uint32_t s1 = 0, s2 = 0;
        const uint32_t l1 = walk_code(e, PBITS, cv, sym, s1);
        uint32_t out = l1 ? (l1 | (s1 << 4)) : 0;
        if (l1 && s1 < 256 && l1 < (uint32_t)PBITS) {
            const uint32_t l2 = walk_code(e >> l1, PBITS - l1, cv, sym, s2);
            if (l2 && s2 < 256) out |= PAIR | ((l1 + l2) << 14) | (s2 << 18);
        }
        lut[e] = out;
    }
    WFENCE();
}

// Global-memory pointers keep their address space through structs and
// calls, so their loads and stores are global (vector memory counter only),
// not flat: a flat access also counts on the LDS counter, and every wait for
// an LDS read would then wait for it too.
#define FR_GLOBAL __attribute__((address_space(1)))
// 16-byte global loads and stores through a native vector type (uint4's
// assignment operators take generic-space references only)
typedef uint32_t fr_v4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 gload16(const FR_GLOBAL void *p)
{
    const fr_v4 x = *reinterpret_cast<const FR_GLOBAL fr_v4 *>(p);
    return make_uint4(x.x, x.y, x.z, x.w);
}
__device__ __forceinline__ void gstore16(FR_GLOBAL void *p, uint4 v)
{
    *reinterpret_cast<FR_GLOBAL fr_v4 *>(p) = fr_v4{v.x, v.y, v.z, v.w};
}
struct Out {
    FR_GLOBAL uint8_t *g; // member's output in HBM
    uint32_t pos;       // bytes produced
    uint32_t flushed;   // bytes written back to HBM (and folded into the CRC)
    uint32_t visible;   // bytes whose HBM stores are known complete
    uint32_t cap;       // ISIZE
    uint32_t raw;       // zero-init CRC register of bytes [0, flushed)
    uint32_t skew;      // ring slot of byte q is (skew + q) & RMASK; skew = output address mod 16 KiB
};

template <uint32_t RING>
struct Geometry {
    static constexpr uint32_t RMASK = RING - 1;
    static constexpr uint32_t FLUSH_AT = RING / 2;       // unflushed bytes that trigger a write-back
    static constexpr uint32_t CHUNK_MAX = FLUSH_AT + 258; // most bytes one write-back can carry
    static constexpr uint32_t SEG = (CHUNK_MAX + 63) / 64; // CRC segment per lane
    // token form: write-back at RING/4 unflushed bytes at a chunk start, and a
    // chunk of tokens stops once it has produced RING/4 - 258 bytes, so a chunk
    // adds under RING/4 and one write-back never carries more than RING/2
    static constexpr uint32_t TOK_FLUSH_AT = RING / 4;
    static constexpr uint32_t TOK_LIMIT = RING / 4 - 258;
};

// Token form of a symbol: a literal is its byte; a match is
// TOK_MATCH | length << 16 | distance.
constexpr uint32_t TOK_MATCH = 1u << 31;
// Matches longer than this are copied by all 64 lanes in the ordered pass.
constexpr uint32_t TOK_SHORT = 16;

// Writes the output of tokens [0, ntok) that start at chunk_pos. First every
// lane takes one token: literals, and matches whose source lies wholly before
// the chunk, are written at once (their sources are final); sources at or
// above `lim` are read from the ring, older ones from HBM, where the write-back
// put them (the chunk's own writes only overwrite ring slots of positions below
// lim). Matches that read this chunk's output, or are long, follow in token
// order with all lanes copying.
template <uint32_t RING>
__device__ void resolve_lane_tokens(uint8_t *ring, uint32_t t_lane, Out &o, uint32_t chunk_pos, uint32_t ntok,
                                    uint32_t lane)
{
    using G = Geometry<RING>;
    const bool live = lane < ntok;
    const uint32_t t = live ? t_lane : 0;
    const bool match = live && (t & TOK_MATCH);
    const uint32_t ln = !live ? 0 : match ? (t >> 16) & 511 : 1;
    uint32_t inc = ln;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if ((int)lane >= d) inc += y;
    }
    const uint32_t end = chunk_pos + (uint32_t)__builtin_amdgcn_readlane((int)inc, 63);
    const uint32_t lim = end > RING ? end - RING : 0;
    const uint32_t dst = chunk_pos + inc - ln;
    const uint32_t dist = t & 0xFFFF;
    const uint32_t src = dst - dist;
    const bool deferred = match && (src + ln > chunk_pos || ln > TOK_SHORT);
    if (__ballot(match && src < lim) != 0 && o.visible < o.flushed) {
        // reads from HBM below: the write-back's stores must have completed
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        o.visible = o.flushed;
    }
    if (live && !match) {
        ring[(o.skew + dst) & G::RMASK] = (uint8_t)t;
    } else if (match && !deferred) {
        for (uint32_t i = 0; i < ln; ++i) {
            const uint32_t p = src + i;
            const uint8_t v = p >= lim ? ring[(o.skew + p) & G::RMASK] : o.g[p];
            ring[(o.skew + dst + i) & G::RMASK] = v;
        }
    }
    WFENCE();
    uint64_t pending = __ballot(deferred);
    while (pending) {
        const uint32_t j = (uint32_t)__ffsll((unsigned long long)pending) - 1;
        pending &= pending - 1;
        const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)dst, (int)j);
        const uint32_t L = (uint32_t)__builtin_amdgcn_readlane((int)ln, (int)j);
        const uint32_t D = (uint32_t)__builtin_amdgcn_readlane((int)dist, (int)j);
        if (D >= L) {
            for (uint32_t off = lane; off < L; off += 64) {
                const uint32_t p = d0 - D + off;
                const uint8_t v = p >= lim ? ring[(o.skew + p) & G::RMASK] : o.g[p];
                ring[(o.skew + d0 + off) & G::RMASK] = v;
            }
        } else {
            // period-D pattern: every source byte precedes d0 and lies in the ring
            uint32_t r = lane % D;
            const uint32_t step = 64u % D;
            for (uint32_t off = lane; off < L; off += 64) {
                ring[(o.skew + d0 + off) & G::RMASK] = ring[(o.skew + d0 - D + r) & G::RMASK];
                r += step;
                if (r >= D) r -= D;
            }
        }
        WFENCE();
    }
}

// Token list held in LDS (one per lane).
template <uint32_t RING>
__device__ void resolve_tokens(uint8_t *ring, const uint32_t *tok, Out &o, uint32_t chunk_pos, uint32_t ntok,
                               uint32_t lane)
{
    WFENCE();
    resolve_lane_tokens<RING>(ring, lane < ntok ? tok[lane] : 0, o, chunk_pos, ntok, lane);
}

// Writes ring bytes [flushed, pos) to HBM and folds them into the CRC: lane l
// takes the l-th of 64 SEG-byte segments counted back from the end (leading
// zeros do not change a zero-initialised CRC register), then the segments are
// combined with x^(8*SEG*(63-l)) and an XOR reduction. The HBM write uses
// 16-byte stores for every aligned 16-byte word inside the range and bytes at
// its edges (neighbouring members own the rest of those words).
template <uint32_t RING>
__device__ void flush(const uint8_t *ring, const uint32_t *crctab, Out &o, uint32_t lane, uint32_t lane_factor,
                      const uint32_t *x2n)
{
    using G = Geometry<RING>;
    const uint32_t c = o.pos - o.flushed;
    if (c == 0) return;
    WFENCE();
    FR_GLOBAL uint8_t *dst = o.g + o.flushed;
    const uint32_t s0 = o.skew + o.flushed; // slot (unmasked) of the first byte
    const uint32_t head = (16u - (uint32_t)((uintptr_t)dst & 15)) & 15;
    if (c < head + 16) {
        for (uint32_t i = lane; i < c; i += 64) dst[i] = ring[(s0 + i) & G::RMASK];
    } else {
        const uint32_t words = (c - head) >> 4;
        const uint32_t tail0 = head + (words << 4);
        if (lane < head) dst[lane] = ring[(s0 + lane) & G::RMASK];
        for (uint32_t w = lane; w < words; w += 64) {
            const uint32_t off = head + (w << 4);
            // an aligned output word sits at a 16-byte aligned ring slot
            const uint4 v = *reinterpret_cast<const uint4 *>(ring + ((s0 + off) & G::RMASK));
            gstore16(dst + off, v);
        }
        for (uint32_t i = tail0 + lane; i < c; i += 64) dst[i] = ring[(s0 + i) & G::RMASK];
    }
    int32_t hi = (int32_t)c - (int32_t)(G::SEG * (63 - lane));
    int32_t lo = hi - (int32_t)G::SEG;
    if (lo < 0) lo = 0;
    uint32_t r = 0;
    for (int32_t k = lo; k < hi; ++k) {
        uint32_t byte = ring[(s0 + (uint32_t)k) & G::RMASK];
        r = crctab[(r ^ byte) & 0xFF] ^ (r >> 8);
    }
    r = fr_mulmod(r, lane_factor);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) r ^= __shfl_xor(r, off, 64);
    uint32_t chunk = uni(r);
    o.raw = fr_mulmod(o.raw, fr_xpow8(x2n, c)) ^ chunk;
    o.flushed = o.pos;
}

// Canonical walk over the low bits of `bits` (codes up to 15 bits, `avail`
// of them present): returns the code length, or 0 with err set when the input
// ends inside the code or no code matches, as decode_litlen's slow path.
__device__ __forceinline__ uint32_t walk_bits(uint32_t bits, uint32_t avail, const uint16_t *count, const uint16_t *sym,
                                              uint32_t bad, uint32_t &out_sym, uint32_t &err)
{
    int code = 0, first = 0, index = 0;
    for (int len = 1; len <= 15; ++len) {
        if ((uint32_t)len > avail) {
            err = ST_TRUNCATED;
            return 0;
        }
        code |= (int)((bits >> (len - 1)) & 1);
        int c = (int)uni(count[len]);
        if (code - first < c) {
            out_sym = uni(sym[index + code - first]);
            return (uint32_t)len;
        }
        index += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    err = bad;
    return 0;
}

// Makes stage bytes [lo, lo + 24) resident (lo a multiple of 8, relative to sbase).
__device__ __forceinline__ void ensure_span(uint32_t *stage, Bits &b, uint32_t lo, uint32_t lane)
{
    const uint32_t hi = lo + 24;
    if (hi <= b.stage_end && lo + STAGE_BYTES >= b.stage_end) return;
    if (b.stage_end != 0 && lo + CHUNK >= b.stage_end && hi <= b.stage_end + CHUNK) {
        load_chunk(stage, b.sbase, b.stage_end, lane);
        b.stage_end += CHUNK;
    } else {
        const uint32_t c0 = lo & ~(CHUNK - 1);
        load_chunk(stage, b.sbase, c0, lane);
        load_chunk(stage, b.sbase, c0 + CHUNK, lane);
        b.stage_end = c0 + 2 * CHUNK;
    }
    WFENCE();
}

__device__ __forceinline__ uint32_t put_lane(uint32_t v, uint32_t at, uint32_t old)
{
    return __lane_id() == at ? v : old;
}

template <uint32_t RING, int TOK>
__device__ void inflate_member(WaveMem<RING> &sm, const uint32_t *crctab, const uint8_t *gin, uint64_t in_off,
                               uint32_t in_len, uint8_t *gout, uint32_t cap, uint32_t lane, uint32_t lane_factor,
                               const uint32_t *x2n, uint32_t &status_out, uint32_t &produced_out, uint32_t &crc_out)
{
    using G = Geometry<RING>;
    Bits b;
    b.buf = 0;
    b.cnt = 0;
    b.pos = 0;
    b.len = in_len;
    b.sbase = gin + (in_off & ~(uint64_t)(CHUNK - 1));
    b.lead = (uint32_t)(in_off & (CHUNK - 1));
    b.stage_end = 0;
    Out o;
    o.g = (FR_GLOBAL uint8_t *)gout;
    o.pos = 0;
    o.flushed = 0;
    o.visible = 0;
    o.cap = cap;
    o.raw = 0;
    o.skew = (uint32_t)((uintptr_t)gout & 16383);
    uint32_t err = ST_OK;
    int tables = 0; // 0 none, 1 fixed, 2 dynamic
    uint32_t last;
    do {
        if (b.cnt < 32) refill(sm.stage, b, lane);
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
            const uint8_t *p = b.sbase + b.lead + b.pos;
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
            const uint8_t *src = b.sbase + b.lead + b.pos;
            for (uint32_t done = 0; done < len; done += 256) {
                if (o.pos - o.flushed >= G::FLUSH_AT) flush<RING>(sm.ring, crctab, o, lane, lane_factor, x2n);
                uint32_t piece = len - done < 256 ? len - done : 256;
                for (uint32_t i = lane; i < piece; i += 64)
                    sm.ring[(o.skew + o.pos + i) & G::RMASK] = src[done + i];
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
                build_pair_lut(sm.plut, sm.lcount, sm.lsym, lane);
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
                if (b.cnt < 3) refill(sm.stage, b, lane);
                uint32_t v = getbits(b, 3, err);
                if (err) break;
                if (lane == 0) sm.lens[clen_order(i)] = (uint8_t)v;
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
                if (b.cnt < 32) refill(sm.stage, b, lane);
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
            build_pair_lut(sm.plut, sm.lcount, sm.lsym, lane);
            build_lut<DBITS>(sm.dlut, sm.dcount, sm.dsym, lane);
            tables = 2;
        }
        if constexpr (TOK == 2) {
            // Lane-speculative decode. Each round, lane l looks up the code
            // that would start l bits past the round's start B (one table read
            // per lane, all at once); the scalar chain then hops from symbol to
            // symbol by reading the lanes it lands on, and a symbol whose
            // distance code would start past lane 63 waits for the next round.
            // Symbols become tokens held one per lane in a register, resolved
            // as in the token form.
            const uint32_t total = b.len * 8;
            uint32_t B = b.pos * 8 - b.cnt; // bits consumed, from the payload start
            uint32_t tokv = 0, ntok = 0, chunk_pos = o.pos;
            bool eob = false;
            while (!eob) {
                B = uni(B);
                ntok = uni(ntok);
                chunk_pos = uni(chunk_pos);
                o.pos = uni(o.pos);
                o.flushed = uni(o.flushed);
                b.stage_end = uni(b.stage_end);
                if (ntok > 48 || o.pos - chunk_pos >= G::TOK_LIMIT) {
                    resolve_lane_tokens<RING>(sm.ring, tokv, o, chunk_pos, ntok, lane);
                    ntok = 0;
                    chunk_pos = o.pos;
                }
                if (ntok == 0 && o.pos - o.flushed >= G::TOK_FLUSH_AT)
                    flush<RING>(sm.ring, crctab, o, lane, lane_factor, x2n);
                // lane l: 32 bits from bit B + l
                const uint32_t r0 = b.lead + (B >> 3);
                ensure_span(sm.stage, b, r0 & ~7u, lane);
                const uint32_t bit = (B & 7) + lane;
                const uint32_t byte = r0 + (bit >> 3);
                const uint32_t w = (byte >> 3) & (STAGE_WORDS / 2 - 1);
                const uint2 x = reinterpret_cast<const uint2 *>(sm.stage)[w];
                const uint2 y = reinterpret_cast<const uint2 *>(sm.stage)[(w + 1) & (STAGE_WORDS / 2 - 1)];
                const uint64_t lo = ((uint64_t)x.y << 32) | x.x, hi = ((uint64_t)y.y << 32) | y.x;
                const uint32_t o8 = (byte & 7) * 8;
                const uint64_t v = o8 ? ((lo >> o8) | (hi << (64 - o8))) : lo;
                const uint32_t fl = (uint32_t)(v >> (bit & 7));
                const uint32_t ev = sm.plut[fl & ((1u << PBITS) - 1)];
                const uint32_t dv = sm.dlut[fl & ((1u << DBITS) - 1)];
                uint32_t p = 0;
                // Fast hops while every code this round can read (a symbol
                // starting below bit 64 takes at most 48 bits with its
                // distance) ends before the payload does, so no truncation
                // check applies: a symbol is taken here only when its codes
                // are in the tables and every check passes. Anything else
                // (a long code, the block's end, an error) leaves the symbol
                // to the general loop below, which decodes it again in full.
                if (B + 128 <= total) {
                    while (p < 64 && ntok < 62 && o.pos - chunk_pos < G::TOK_LIMIT) {
                        const uint32_t e = uni(__builtin_amdgcn_readlane(ev, p));
                        const uint32_t L = e & 15, sym = (e >> 4) & 511;
                        if (L == 0) break;
                        if (sym < 256) {
                            if ((e & PAIR) && o.pos + 2 <= o.cap) {
                                tokv = put_lane((e >> 4) & 255, ntok, tokv);
                                tokv = put_lane((e >> 18) & 255, ntok + 1, tokv);
                                ntok += 2;
                                o.pos += 2;
                                p += (e >> 14) & 15;
                                continue;
                            }
                            if (o.pos >= o.cap) break;
                            tokv = put_lane(sym, ntok, tokv);
                            ++ntok;
                            ++o.pos;
                            p += L;
                            continue;
                        }
                        const uint32_t lc = sym - 257;
                        if (lc >= 29) break;
                        const uint32_t f = uni(__builtin_amdgcn_readlane(fl, p));
                        const uint32_t lx = len_extra(lc);
                        const uint32_t len = len_base(lc) + ((f >> L) & ((1u << lx) - 1));
                        const uint32_t q = p + L + lx;
                        if (q >= 64) break;
                        const uint32_t ed = uni(__builtin_amdgcn_readlane(dv, q));
                        const uint32_t Ld = ed & 15, ds = ed >> 4;
                        if (Ld == 0 || ds >= 30) break;
                        const uint32_t g = uni(__builtin_amdgcn_readlane(fl, q));
                        const uint32_t dx = dist_extra(ds);
                        const uint32_t dist = dist_base(ds) + ((g >> Ld) & ((1u << dx) - 1));
                        if (dist > o.pos || o.pos + len > o.cap) break;
                        tokv = put_lane(TOK_MATCH | (len << 16) | dist, ntok, tokv);
                        ++ntok;
                        o.pos += len;
                        p = q + Ld + dx;
                    }
                }
                while (p < 64 && ntok < 62 && o.pos - chunk_pos < G::TOK_LIMIT) {
                    const uint32_t e = uni(__builtin_amdgcn_readlane(ev, p));
                    const uint32_t f = uni(__builtin_amdgcn_readlane(fl, p));
                    uint32_t L = e & 15, sym = (e >> 4) & 511;
                    if (L) {
                        if (B + p + L > total) {
                            err = ST_TRUNCATED;
                            break;
                        }
                    } else {
                        L = walk_bits(f, total - (B + p), sm.lcount, sm.lsym, ST_BAD_CODE, sym, err);
                        if (err) break;
                    }
                    if (sym < 256) {
                        const uint32_t tb = (e >> 14) & 15;
                        if ((e & PAIR) && B + p + tb <= total && o.pos + 2 <= o.cap) {
                            tokv = put_lane((e >> 4) & 255, ntok, tokv);
                            tokv = put_lane((e >> 18) & 255, ntok + 1, tokv);
                            ntok += 2;
                            o.pos += 2;
                            p += tb;
                            continue;
                        }
                        if (o.pos >= o.cap) {
                            err = ST_OUTPUT_OVERFLOW;
                            break;
                        }
                        tokv = put_lane(sym, ntok, tokv);
                        ++ntok;
                        ++o.pos;
                        p += L;
                        continue;
                    }
                    if (sym == 256) {
                        p += L;
                        eob = true;
                        break;
                    }
                    const uint32_t lc = sym - 257;
                    if (lc >= 29) {
                        err = ST_BAD_CODE;
                        break;
                    }
                    // length extra bits follow the code in this lane's 32 bits
                    const uint32_t lx = len_extra(lc);
                    if (B + p + L + lx > total) {
                        err = ST_TRUNCATED;
                        break;
                    }
                    const uint32_t len = len_base(lc) + ((f >> L) & ((1u << lx) - 1));
                    const uint32_t q = p + L + lx;
                    if (q >= 64) break; // distance code past the lanes: next round
                    const uint32_t ed = uni(__builtin_amdgcn_readlane(dv, q));
                    const uint32_t g = uni(__builtin_amdgcn_readlane(fl, q));
                    uint32_t Ld = ed & 15, ds = ed >> 4;
                    if (Ld) {
                        if (B + q + Ld > total) {
                            err = ST_TRUNCATED;
                            break;
                        }
                    } else {
                        Ld = walk_bits(g, total - (B + q), sm.dcount, sm.dsym, ST_BAD_DISTANCE, ds, err);
                        if (err) break;
                    }
                    if (ds >= 30) {
                        err = ST_BAD_DISTANCE;
                        break;
                    }
                    const uint32_t dx = dist_extra(ds);
                    if (B + q + Ld + dx > total) {
                        err = ST_TRUNCATED;
                        break;
                    }
                    const uint32_t dist = dist_base(ds) + ((g >> Ld) & ((1u << dx) - 1));
                    if (dist > o.pos) {
                        err = ST_BAD_DISTANCE;
                        break;
                    }
                    if (o.pos + len > o.cap) {
                        err = ST_OUTPUT_OVERFLOW;
                        break;
                    }
                    tokv = put_lane(TOK_MATCH | (len << 16) | dist, ntok, tokv);
                    ++ntok;
                    o.pos += len;
                    p = q + Ld + dx;
                }
                if (err) break;
                B += p;
            }
            if (err) break;
            if (ntok) resolve_lane_tokens<RING>(sm.ring, tokv, o, chunk_pos, ntok, lane);
            // give the position back to the bit reader for the next block header
            b.buf = 0;
            b.cnt = 0;
            b.pos = B >> 3;
            if (B & 7) {
                refill(sm.stage, b, lane);
                (void)getbits(b, B & 7, err);
            }
            continue;
        }
        if constexpr (TOK == 1) {
            // decode into tokens; resolve them 64 (or TOK_LIMIT bytes) at a time
            uint32_t ntok = 0, chunk_pos = o.pos;
            for (;;) {
                b.buf = uni64(b.buf);
                b.cnt = uni(b.cnt);
                b.pos = uni(b.pos);
                b.stage_end = uni(b.stage_end);
                o.pos = uni(o.pos);
                o.flushed = uni(o.flushed);
                ntok = uni(ntok);
                chunk_pos = uni(chunk_pos);
                if (ntok >= 62 || o.pos - chunk_pos >= G::TOK_LIMIT) {
                    resolve_tokens<RING>(sm.ring, sm.tok, o, chunk_pos, ntok, lane);
                    ntok = 0;
                    chunk_pos = o.pos;
                }
                if (ntok == 0 && o.pos - o.flushed >= G::TOK_FLUSH_AT)
                    flush<RING>(sm.ring, crctab, o, lane, lane_factor, x2n);
                if (b.cnt < 32) refill(sm.stage, b, lane);
                const uint32_t e = uni(sm.plut[b.buf & ((1u << PBITS) - 1)]);
                if ((e & PAIR) && ((e >> 14) & 15) <= b.cnt && o.pos + 2 <= o.cap) {
                    const uint32_t odd = lane & 1;
                    sm.tok[ntok + odd] = odd ? ((e >> 18) & 255) : ((e >> 4) & 255);
                    const uint32_t tb = (e >> 14) & 15;
                    b.buf >>= tb;
                    b.cnt -= tb;
                    o.pos += 2;
                    ntok += 2;
                    continue;
                }
                int sym = decode_litlen(b, e, sm.lcount, sm.lsym, err);
                if (err) break;
                if (sym < 256) {
                    if (o.pos >= o.cap) {
                        err = ST_OUTPUT_OVERFLOW;
                        break;
                    }
                    sm.tok[ntok] = (uint32_t)sym;
                    ++ntok;
                    ++o.pos;
                    continue;
                }
                if (sym == 256) break;
                const uint32_t lc = (uint32_t)sym - 257;
                if (lc >= 29) {
                    err = ST_BAD_CODE;
                    break;
                }
                const uint32_t len = len_base(lc) + getbits(b, len_extra(lc), err);
                if (err) break;
                if (b.cnt < 28) refill(sm.stage, b, lane);
                int ds = decode_sym<DBITS>(b, sm.dlut, sm.dcount, sm.dsym, ST_BAD_DISTANCE, err);
                if (err) break;
                if (ds >= 30) {
                    err = ST_BAD_DISTANCE;
                    break;
                }
                const uint32_t dist = dist_base((uint32_t)ds) + getbits(b, dist_extra((uint32_t)ds), err);
                if (err) break;
                if (dist > o.pos) {
                    err = ST_BAD_DISTANCE;
                    break;
                }
                if (o.pos + len > o.cap) {
                    err = ST_OUTPUT_OVERFLOW;
                    break;
                }
                sm.tok[ntok] = TOK_MATCH | (len << 16) | dist;
                ++ntok;
                o.pos += len;
            }
            if (err) break;
            if (ntok) resolve_tokens<RING>(sm.ring, sm.tok, o, chunk_pos, ntok, lane);
            continue;
        }
        // literal / length / distance codes until end-of-block
        for (;;) {
            // The state is the same in every lane; say so, so that it stays in
            // SGPRs and the loop runs on the scalar unit.
            b.buf = uni64(b.buf);
            b.cnt = uni(b.cnt);
            b.pos = uni(b.pos);
            b.stage_end = uni(b.stage_end);
            o.pos = uni(o.pos);
            o.flushed = uni(o.flushed);
            if (o.pos - o.flushed >= G::FLUSH_AT) flush<RING>(sm.ring, crctab, o, lane, lane_factor, x2n);
            if (b.cnt < 32) refill(sm.stage, b, lane);
            const uint32_t e = uni(sm.plut[b.buf & ((1u << PBITS) - 1)]);
            if ((e & PAIR) && ((e >> 14) & 15) <= b.cnt && o.pos + 2 <= o.cap) {
                // two literals: even lanes store the first, odd lanes the second
                // (every lane stores, so no lane mask is set up)
                const uint32_t odd = lane & 1;
                sm.ring[(o.skew + o.pos + odd) & G::RMASK] = (uint8_t)(odd ? (e >> 18) : (e >> 4));
                const uint32_t t = (e >> 14) & 15;
                b.buf >>= t;
                b.cnt -= t;
                o.pos += 2;
                continue;
            }
            int sym = decode_litlen(b, e, sm.lcount, sm.lsym, err);
            if (err) break;
            if (sym < 256) {
                if (o.pos >= o.cap) {
                    err = ST_OUTPUT_OVERFLOW;
                    break;
                }
                sm.ring[(o.skew + o.pos) & G::RMASK] = (uint8_t)sym;
                ++o.pos;
                continue;
            }
            if (sym == 256) break;
            const uint32_t lc = (uint32_t)sym - 257;
            if (lc >= 29) {
                err = ST_BAD_CODE;
                break;
            }
            uint32_t len = len_base(lc) + getbits(b, len_extra(lc), err);
            if (err) break;
            if (b.cnt < 28) refill(sm.stage, b, lane);
            int ds = decode_sym<DBITS>(b, sm.dlut, sm.dcount, sm.dsym, ST_BAD_DISTANCE, err);
            if (err) break;
            if (ds >= 30) {
                err = ST_BAD_DISTANCE;
                break;
            }
            uint32_t dist = dist_base((uint32_t)ds) + getbits(b, dist_extra((uint32_t)ds), err);
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
            const uint32_t d0 = o.skew + o.pos;
            if (dist + len <= RING) {
                const uint32_t s0 = d0 - dist;
                if (dist >= len) {
                    for (uint32_t off = lane; off < len; off += 64)
                        sm.ring[(d0 + off) & G::RMASK] = sm.ring[(s0 + off) & G::RMASK];
                } else {
                    // period-`dist` pattern: every source byte precedes o.pos
                    uint32_t r = lane % dist;
                    uint32_t step = 64u % dist;
                    for (uint32_t off = lane; off < len; off += 64) {
                        sm.ring[(d0 + off) & G::RMASK] = sm.ring[(s0 + r) & G::RMASK];
                        r += step;
                        if (r >= dist) r -= dist;
                    }
                }
            } else {
                // Beyond the ring: these bytes were written back to HBM by this
                // wave (dist > RING - 258 puts them below `flushed`, and there
                // is no overlap since dist > len). Make those stores complete
                // before reading them back.
                if (o.visible < o.flushed) {
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                    o.visible = o.flushed;
                }
                const FR_GLOBAL uint8_t *src = o.g + (o.pos - dist);
                for (uint32_t off = lane; off < len; off += 64) sm.ring[(d0 + off) & G::RMASK] = src[off];
            }
            WFENCE();
            o.pos += len;
        }
        if (err) break;
    } while (!last);

    if (err == ST_OK) {
        flush<RING>(sm.ring, crctab, o, lane, lane_factor, x2n);
        crc_out = o.raw ^ fr_mulmod(0xFFFFFFFFu, fr_xpow8(x2n, o.pos)) ^ 0xFFFFFFFFu;
    } else {
        crc_out = 0;
    }
    status_out = err;
    produced_out = o.pos;
}

// NW waves per workgroup, each inflating its own members; the CRC table is
// shared by the workgroup. Members come from the work-queue counter when one is
// given, else member (block, wave) of a grid-stride walk.
template <uint32_t RING, uint32_t NW, int TOK>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void fr_inflate_kernel(InflateArgs a)
{
    static_assert(RING >= 4096 && RING <= 32768, "ring size");
    __shared__ uint32_t crctab[256];
    __shared__ __attribute__((aligned(16))) WaveMem<RING> waves[NW];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < 256; i += 64 * NW) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        crctab[i] = c;
    }
    __syncthreads();
    WaveMem<RING> &sm = waves[wave];
    const uint32_t lane_factor = fr_xpow8(a.x2n, (uint64_t)Geometry<RING>::SEG * (63 - lane));
    uint32_t iter = 0;
    for (;;) {
        uint32_t m;
        if (a.queue) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(a.queue, 1u);
            m = uni(t);
        } else {
            m = (blockIdx.x + iter * gridDim.x) * NW + wave;
        }
        ++iter;
        if (m >= a.n) break;
        if (a.order) m = a.order[m];
        uint32_t st, produced, crc;
        // member parameters are wave-uniform: keep them (and all the decode
        // derives from them) scalar; pointers stay typed, only offsets are made uniform
        inflate_member<RING, TOK>(sm, crctab, a.in, uni64(a.in_off[m]), uni(a.in_len[m]), a.out + uni64(a.out_off[m]),
                             uni(a.out_cap[m]), lane, lane_factor, a.x2n, st, produced, crc);
        if (lane == 0) {
            a.status[m] = st;
            a.produced[m] = produced;
            a.crc[m] = crc;
        }
    }
}

// ---------------------------------------------------------------------------
// Two-phase form: the Huffman decode runs one lane per member, the copies one
// wave per member.
//
// Phase 1 (fr_inflate_lanes_kernel): lane t of the grid owns member t. Its
// decode is per lane -- canonical limits in VGPRs, symbol lists sorted by
// (code length, value) in lane-minor LDS -- so 64 members advance one symbol
// per step on the vector unit. Literals and stored bytes go straight to their
// output offsets; each match becomes a 32-bit record in a scratch list:
// bits 23-31 the count of literal bytes since the previous record's span
// (REC_SKIP: a record that only covers REC_SKIP literal bytes), bits 15-22
// length - 3, bits 0-14 distance - 1. Status, produced bytes and the record
// count are final after this phase; the CRC is left to phase 2.
// Phase 2 (fr_inflate_resolve_kernel): one wave per member walks its records
// 64 at a time. The output span they cover (literals already in place) is
// read into the LDS ring, the matches are copied as in the token form, and
// the ring is written back with the CRC folded in.

constexpr uint32_t REC_SKIP = 511;
// Decode steps per pass of the lanes kernel. On the 20M-record BAM: 3, 5 and
// 6 within 0.3 ms of 4 (36.0 ms); 8 slower (37.9 ms: lanes that use up their
// staged input mid-pass skip the rest of it).
constexpr uint32_t LSTEPS = 4;
constexpr uint32_t LQ = 14; // packed limits for code lengths 1..14

// First record slot of member m. Each record covers at least 3 output bytes,
// so a member of ISIZE c needs at most c / 3 slots, below its share here.
__device__ __forceinline__ uint64_t rec_base(uint64_t out_off, uint32_t m) { return (out_off * 11) / 32 + 2ull * m; }

// Canonical code in limit form. With v the next 15 bits read MSB-first, the
// next code has length L exactly when lim[L-1] <= v < lim[L] (lim[0] = 0,
// lim[k] = (first code of length k + count[k]) << (15 - k)), and its symbol is
// entry offs[L] + ((v - lim[L-1]) >> (15 - L)) of the sorted list. q[k-1]
// packs lim[k] << 16 | (k + 1) << 12 | offs[k + 1] for k = 1..14.
struct LaneCode {
    uint32_t q[LQ];
    uint32_t lim15;
};

// Builds `c` from counts cnt(1..15) of n symbols; returns puff's construct
// value: 0 complete (or no codes at all), > 0 incomplete, < 0 over-subscribed.
template <typename Cnt>
__device__ __forceinline__ int lane_code(LaneCode &c, uint32_t n, Cnt cnt)
{
    uint32_t code = 0, offs = 0;
    int left = 1;
#pragma unroll
    for (int k = 1; k <= 15; ++k) {
        const uint32_t ck = cnt(k);
        left = (left << 1) - (int)ck;
        const uint32_t lim = (code + ck) << (15 - k);
        if (k < 15)
            c.q[k - 1] = (lim << 16) | ((uint32_t)(k + 1) << 12) | (offs + ck);
        else
            c.lim15 = lim;
        code = (code + ck) << 1;
        offs += ck;
    }
    if (offs == 0) return 0;
    (void)n;
    return left;
}

// Sorted-list index of the code at the front of v (15 bits, MSB first); L = 0
// when v starts with no code of this table.
// MAXL: the longest code the table can hold (15, or 7 for the code-length
// code, whose limits for lengths 8..15 all equal lim[7]).
template <int MAXL = 15>
__device__ __forceinline__ uint32_t lane_find(const LaneCode &c, uint32_t v, uint32_t &L)
{
    // q[] increases with k, so the entry that applies is the largest of
    // A = {1 << 12 (length 1), q[0], ..., q[MAXL-2]} at or below v16
    const uint32_t v16 = (v << 16) | 0xFFFFu;
    uint32_t sel;
    if constexpr (MAXL == 15) {
        // binary search over A padded to 16 entries with ~0: four compares,
        // each halving the candidates with selects
        const bool c1 = v16 >= c.q[7];
        const uint32_t B0 = c1 ? c.q[7] : (1u << 12), B1 = c1 ? c.q[8] : c.q[0], B2 = c1 ? c.q[9] : c.q[1],
                       B3 = c1 ? c.q[10] : c.q[2], B4 = c1 ? c.q[11] : c.q[3], B5 = c1 ? c.q[12] : c.q[4],
                       B6 = c1 ? c.q[13] : c.q[5], B7 = c1 ? ~0u : c.q[6];
        const bool c2 = v16 >= B4;
        const uint32_t C0 = c2 ? B4 : B0, C1 = c2 ? B5 : B1, C2 = c2 ? B6 : B2, C3 = c2 ? B7 : B3;
        const bool c3 = v16 >= C2;
        const uint32_t D0 = c3 ? C2 : C0, D1 = c3 ? C3 : C1;
        sel = v16 >= D1 ? D1 : D0;
    } else {
        sel = 1u << 12;
#pragma unroll
        for (int k = 0; k < MAXL - 1; ++k) sel = v16 >= c.q[k] ? c.q[k] : sel;
    }
    const uint32_t len = (sel >> 12) & 15;
    const uint32_t lim = MAXL == 15 ? c.lim15 : c.q[MAXL - 1] >> 16;
    L = v < lim ? len : 0;
    return (sel & 0xFFFu) + ((v - (sel >> 16)) >> (15 - len));
}

// Per-lane LSB-first bit reader. Input is staged in two 32-byte buffers: the
// decode takes words only from A, a queue whose next word is always A0.x;
// B holds the next 32 bytes. B is read (moved into A) and refetched only at
// points where the whole wave does so -- every few decode steps, or in the
// header path -- because a wait for a load is a wave-wide counter: a buffer
// that some lane reads every step must never have a load in flight.
struct LaneBits {
    uint64_t bb;      // next bits, LSB first (may run past the payload)
    uint32_t bc;      // bits in bb
    uint32_t avail;   // payload bits not yet consumed
    uint32_t k;       // words of A used (8: A used up)
    uint4 A0, A1, B0, B1;
    const uint4 *ip;  // next 32 bytes to fetch
};

// Drops the front word of A (a dynamic index into A would put the reader in
// scratch memory).
#define LB_SHIFT(b)                                                                                       \
    do {                                                                                                  \
        (b).A0.x = (b).A0.y;                                                                              \
        (b).A0.y = (b).A0.z;                                                                              \
        (b).A0.z = (b).A0.w;                                                                              \
        (b).A0.w = (b).A1.x;                                                                              \
        (b).A1.x = (b).A1.y;                                                                              \
        (b).A1.y = (b).A1.z;                                                                              \
        (b).A1.z = (b).A1.w;                                                                              \
    } while (0)

// Moves B into A and refetches B. The empty statement (with a memory
// clobber) keeps the loads after the moves, so they land in B's own
// registers. Macros, not functions: compiled on their own, the conditional
// block moves would go through memory.
#define LB_SWAP(b)                                                                                        \
    do {                                                                                                  \
        (b).A0 = (b).B0;                                                                                  \
        (b).A1 = (b).B1;                                                                                  \
        asm volatile("" : "+v"((b).A0.x), "+v"((b).A0.y), "+v"((b).A0.z), "+v"((b).A0.w), "+v"((b).A1.x),   \
                     "+v"((b).A1.y), "+v"((b).A1.z), "+v"((b).A1.w)::"memory");                          \
        (b).B0 = (b).ip[0];                                                                               \
        (b).B1 = (b).ip[1];                                                                               \
        (b).ip += 2;                                                                                      \
        (b).k = 0;                                                                                        \
    } while (0)

// Tops bb up from A when it holds under 32 bits and A has words left.
#define LB_REFILL(b)                                                                                      \
    do {                                                                                                  \
        if ((b).bc < 32 && (b).k < 8) {                                                                   \
            (b).bb |= (uint64_t)(b).A0.x << (b).bc;                                                       \
            (b).bc += 32;                                                                                 \
            ++(b).k;                                                                                      \
            LB_SHIFT(b);                                                                                  \
        }                                                                                                 \
    } while (0)

// The same as selects, for the decode steps (no branch: most steps of some
// lane refill, so the branch would be taken anyway).
#define LB_REFILL_SEL(b)                                                                                  \
    do {                                                                                                  \
        const bool rf_ = (b).bc < 32 && (b).k < 8;                                                        \
        (b).bb |= (uint64_t)(rf_ ? (b).A0.x : 0u) << ((b).bc & 31);                                       \
        (b).bc += rf_ ? 32u : 0u;                                                                         \
        (b).k += rf_ ? 1u : 0u;                                                                           \
        (b).A0.x = rf_ ? (b).A0.y : (b).A0.x;                                                             \
        (b).A0.y = rf_ ? (b).A0.z : (b).A0.y;                                                             \
        (b).A0.z = rf_ ? (b).A0.w : (b).A0.z;                                                             \
        (b).A0.w = rf_ ? (b).A1.x : (b).A0.w;                                                             \
        (b).A1.x = rf_ ? (b).A1.y : (b).A1.x;                                                             \
        (b).A1.y = rf_ ? (b).A1.z : (b).A1.y;                                                             \
        (b).A1.z = rf_ ? (b).A1.w : (b).A1.z;                                                             \
    } while (0)

// Same, moving B in first when A is used up (waits for B's load).
#define LB_REFILL_SYNC(b)                                                                                 \
    do {                                                                                                  \
        if ((b).bc < 32) {                                                                                \
            if ((b).k == 8) LB_SWAP(b);                                                                   \
            LB_REFILL(b);                                                                                 \
        }                                                                                                 \
    } while (0)

// n <= 16 bits; bb must hold at least n bits.
__device__ __forceinline__ uint32_t lb_bits(LaneBits &b, uint32_t n, uint32_t &err)
{
    if (n > b.avail) {
        err = ST_TRUNCATED;
        return 0;
    }
    const uint32_t v = (uint32_t)b.bb & ((1u << n) - 1);
    b.bb >>= n;
    b.bc -= n;
    b.avail -= n;
    return v;
}

// Next code of table c: its sorted-list index, or err (puff's order: input
// that ends inside a code is truncated, else no code after 15 bits is `bad`).
template <int MAXL = 15>
__device__ __forceinline__ uint32_t lb_decode(LaneBits &b, const LaneCode &c, uint32_t bad, uint32_t &err)
{
    const uint32_t v = __builtin_bitreverse32((uint32_t)b.bb) >> 17;
    uint32_t L;
    const uint32_t idx = lane_find<MAXL>(c, v, L);
    if (L == 0 || L > b.avail) {
        err = L == 0 && b.avail >= 15 ? bad : ST_TRUNCATED;
        return 0;
    }
    b.bb >>= L;
    b.bc -= L;
    b.avail -= L;
    return idx;
}

// Counts per code length (0..15) in 16-bit (F16) or 8-bit (F8) fields of
// named 64-bit words: the length is per lane, and an array indexed by it
// would live in scratch memory.
struct F16 {
    uint64_t w0, w1, w2, w3;
};
struct F8 {
    uint64_t w0, w1;
};
__device__ __forceinline__ uint64_t sel_mask(bool c) { return 0ull - (uint64_t)c; }
__device__ __forceinline__ uint32_t f16_get(const F16 &f, uint32_t L)
{
    const uint32_t g = L >> 2;
    const uint64_t w = (f.w0 & sel_mask(g == 0)) | (f.w1 & sel_mask(g == 1)) | (f.w2 & sel_mask(g == 2)) |
                       (f.w3 & sel_mask(g == 3));
    return (uint32_t)(w >> ((L & 3) * 16)) & 0xFFFFu;
}
__device__ __forceinline__ void f16_add(F16 &f, uint32_t L, uint32_t v)
{
    const uint64_t d = (uint64_t)v << ((L & 3) * 16);
    const uint32_t g = L >> 2;
    f.w0 += d & sel_mask(g == 0);
    f.w1 += d & sel_mask(g == 1);
    f.w2 += d & sel_mask(g == 2);
    f.w3 += d & sel_mask(g == 3);
}
__device__ __forceinline__ uint32_t f8_get(const F8 &f, uint32_t L)
{
    const uint64_t m = sel_mask(L < 8);
    return (uint32_t)(((f.w0 & m) | (f.w1 & ~m)) >> ((L & 7) * 8)) & 0xFFu;
}
__device__ __forceinline__ void f8_add(F8 &f, uint32_t L, uint32_t v)
{
    const uint64_t d = (uint64_t)v << ((L & 7) * 8);
    const uint64_t m = sel_mask(L < 8);
    f.w0 += d & m;
    f.w1 += d & ~m;
}
__device__ __forceinline__ void lb_copy(LaneBits &d, const LaneBits &s)
{
    d.bb = s.bb;
    d.bc = s.bc;
    d.avail = s.avail;
    d.k = s.k;
    d.A0 = s.A0;
    d.A1 = s.A1;
    d.B0 = s.B0;
    d.B1 = s.B1;
    d.ip = s.ip;
}

enum : uint32_t { LS_HEADER = 0, LS_CODES = 1, LS_DONE = 2, LS_LENS = 3, LS_TABLES = 4 };

// Counts per code length of the ndist distance code lengths held 4 bits
// each in d0..d3.
__device__ __forceinline__ F8 dist_counts(uint32_t d0, uint32_t d1, uint32_t d2, uint32_t d3, uint32_t ndist)
{
    F8 c = {0, 0};
#pragma unroll
    for (uint32_t j = 0; j < 30; ++j) {
        const uint32_t dn = j < 8 ? d0 : j < 16 ? d1 : j < 24 ? d2 : d3;
        if (j < ndist) f8_add(c, (dn >> (4 * (j & 7))) & 15, 1);
    }
    return c;
}

template <uint32_t NW>
__global__ __launch_bounds__(64 * NW) void fr_inflate_lanes_kernel(InflateArgs a)
{
    constexpr uint32_t LANES = 64 * NW;
    // Per lane, in lane-minor 32-bit words (word w of lane t at [w * LANES + t]):
    // the literal/length symbol list as bytes (symbol & 255; 72 words), one bit
    // per entry set when the symbol is 256 or above (9 words), the distance
    // list as bytes (8 words). 356 bytes per lane: 7 waves per CU.
    __shared__ uint32_t tab_all[(72 + 9 + 8) * LANES];
    const uint32_t t = threadIdx.x;
    const uint32_t slot = blockIdx.x * LANES + t;
    if (slot >= a.n) return;
    const uint32_t m = a.order ? a.order[slot] : slot;
    uint32_t *lw = tab_all + t;
    uint32_t *hw = tab_all + 72 * LANES + t;
    uint32_t *dw = tab_all + 81 * LANES + t;
#define LANE_BYTE(base, e) (reinterpret_cast<uint8_t *>((base) + ((e) >> 2) * LANES)[(e) & 3])
    const uint64_t in_off = a.in_off[m];
    const uint32_t in_len = a.in_len[m];
    const uint64_t out_off = a.out_off[m];
    const uint32_t cap = a.out_cap[m];
    uint8_t *out = a.out + out_off;
    uint32_t *rec = a.rec + rec_base(out_off, m);

    LaneBits b;
    {
        // the input base is 256-byte aligned; pointer arithmetic (not integer
        // casts) keeps the loads in the global address space
        const uint4 *blk = reinterpret_cast<const uint4 *>(a.in + (in_off & ~15ull));
        b.A0 = blk[0];
        b.A1 = blk[1];
        b.B0 = blk[2];
        b.B1 = blk[3];
        b.ip = blk + 4;
        const uint32_t byte = (uint32_t)(in_off & 15);
#pragma unroll
        for (uint32_t i = 0; i < 3; ++i)
            if (i < (byte >> 2)) LB_SHIFT(b);
        b.k = byte >> 2;
        b.bb = 0;
        b.bc = 0;
        b.avail = in_len * 8;
        LB_REFILL(b);
        b.bb >>= (byte & 3) * 8;
        b.bc -= (byte & 3) * 8;
    }
    LaneCode lc, dc;
    uint32_t state = LS_HEADER, last = 0, err = ST_OK, pos = 0, nrec = 0, mark = 0;

    // Output stores are held in registers and issued together in the next
    // pass, right after B's loads: a wait for B (one pass later) then finds
    // only stores as old as those loads in flight (the memory counter covers
    // stores too, in issue order).
    // Literals are stored as whole 4-byte words of the output ("windows",
    // numbered from the aligned word at or below the member's first byte):
    // the bytes of a word that are not literals belong to matches, which the
    // resolve phase overwrites. Bytes outside [lo, cap) of the member (lo:
    // the member start, or the end of its last stored block, whose bytes are
    // written directly) are never written this way.
    constexpr uint32_t NONE = ~0u;
    const uint32_t h = (uint32_t)(out_off & 3);
    uint8_t *const out_w = out - h; // word-aligned: byte r of the windows is out_w[r]
    uint32_t lo = h;                 // first window byte a word store may cover
    uint32_t wcur = NONE, wv = 0;    // the open window and its bytes
    uint32_t ls_i[LSTEPS], ls_v[LSTEPS]; // windows closed in step j of the pass
    uint32_t rs_i[LSTEPS], rs_v[LSTEPS]; // record written in step j of the pass
#pragma unroll
    for (uint32_t j = 0; j < LSTEPS; ++j) {
        ls_i[j] = NONE;
        ls_v[j] = 0;
        rs_i[j] = NONE;
        rs_v[j] = 0;
    }
    // Window bytes r in [lo, min(end, h + cap)), as one word when all four are.
    auto put_window = [&](uint32_t wi, uint32_t v, uint32_t end) __attribute__((always_inline)) {
        const uint32_t r0 = 4 * wi;
        if (r0 >= lo && r0 + 4 <= end && r0 - h + 4 <= cap) {
            *reinterpret_cast<uint32_t *>(out_w + r0) = v;
        } else {
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j)
                if (r0 + j >= lo && r0 + j < end && r0 + j - h < cap) out_w[r0 + j] = (uint8_t)(v >> (8 * j));
        }
    };
    auto flush_pass = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (uint32_t j = 0; j < LSTEPS; ++j) {
            if (ls_i[j] != NONE) put_window(ls_i[j], ls_v[j], NONE);
            if (rs_i[j] != NONE) rec[rs_i[j]] = rs_v[j];
            ls_i[j] = NONE;
            rs_i[j] = NONE;
        }
    };
    // A dynamic block's code lengths (LS_LENS) take two passes over the same
    // input, lbit bits into the payload: pass 0 counts the literal/length
    // lengths per code length (cll) and keeps the distance lengths, 4 bits
    // each, in dn0..dn3; pass 1 turns the counts into list entries and
    // places the literal/length symbols. There is no room in LDS to keep all
    // the lengths between the passes.
    uint32_t pass = 0, nlen = 0, ndist = 0, total = 0, index = 0, prev = 0, len256 = 0, lbit = 0;
    uint32_t dn0 = 0, dn1 = 0, dn2 = 0, dn3 = 0;
    F16 cll = {0, 0, 0, 0};
    // Puts the reader `bit` bits into the payload. It waits for its loads:
    // rare (twice per dynamic block), and no later step then waits for them.
    auto lb_seek = [&](uint32_t bit) __attribute__((always_inline)) {
        const uint64_t at = in_off + (bit >> 3);
        const uint4 *blk = reinterpret_cast<const uint4 *>(a.in + (at & ~15ull));
        b.A0 = blk[0];
        b.A1 = blk[1];
        b.B0 = blk[2];
        b.B1 = blk[3];
        b.ip = blk + 4;
        __builtin_amdgcn_s_waitcnt(0);
        const uint32_t byte = (uint32_t)(at & 15);
#pragma unroll
        for (uint32_t i = 0; i < 3; ++i)
            if (i < (byte >> 2)) LB_SHIFT(b);
        b.k = byte >> 2;
        b.bb = 0;
        b.bc = 0;
        b.avail = in_len * 8 - bit;
        LB_REFILL(b);
        const uint32_t drop = (byte & 3) * 8 + (bit & 7);
        b.bb >>= drop;
        b.bc -= drop;
    };
    // the member's parameters have landed: otherwise their first use inside
    // the loop would leave a wait on the memory counter in every step
    __builtin_amdgcn_s_waitcnt(0);
    // Each pass of the loop: B moves into A where A is used up (the same
    // point in every lane, so no decode step waits for memory), the stores of
    // the last pass, the block header for lanes at one, then four decode
    // steps.
    while (state != LS_DONE) {
        if (b.k == 8) LB_SWAP(b);
        asm volatile("" ::: "memory"); // the stores stay after the loads
        flush_pass();
        if (state == LS_HEADER) {
                // block header
                LB_REFILL_SYNC(b);
                last = lb_bits(b, 1, err);
                const uint32_t type = lb_bits(b, 2, err);
                if (err != ST_OK) {
                    state = LS_DONE;
                    continue;
                }
                if (type == 0) {
                    // stored: to a byte boundary, then LEN, NLEN and the bytes
                    const uint32_t skip = (in_len * 8 - b.avail) & 7;
                    (void)lb_bits(b, skip ? 8 - skip : 0, err);
                    if (b.avail < 32) {
                        err = ST_TRUNCATED;
                    } else {
                        LB_REFILL_SYNC(b);
                        const uint32_t len = lb_bits(b, 16, err);
                        LB_REFILL_SYNC(b);
                        const uint32_t ncomp = lb_bits(b, 16, err);
                        if (len != (~ncomp & 0xFFFFu))
                            err = ST_STORED_LEN;
                        else if (len * 8 > b.avail)
                            err = ST_TRUNCATED;
                        else if (pos + len > cap)
                            err = ST_OUTPUT_OVERFLOW;
                        else {
                            // the open window's bytes before pos, then the
                            // stored bytes; later windows start after them
                            if (wcur != NONE) put_window(wcur, wv, pos + h);
                            wcur = NONE;
                            wv = 0;
                            // four bytes per step while they last (the reader
                            // is at a byte boundary, and len * 8 <= avail)
                            uint32_t i = 0;
                            for (; i + 4 <= len; i += 4) {
                                LB_REFILL_SYNC(b);
                                const uint32_t v = (uint32_t)b.bb;
                                b.bb >>= 32;
                                b.bc -= 32;
                                b.avail -= 32;
                                out[pos] = (uint8_t)v;
                                out[pos + 1] = (uint8_t)(v >> 8);
                                out[pos + 2] = (uint8_t)(v >> 16);
                                out[pos + 3] = (uint8_t)(v >> 24);
                                pos += 4;
                            }
                            for (; i < len; ++i) {
                                LB_REFILL_SYNC(b);
                                out[pos++] = (uint8_t)lb_bits(b, 8, err);
                            }
                            lo = pos + h;
                        }
                    }
                    state = err != ST_OK || last ? LS_DONE : LS_HEADER;
                    continue;
                }
                if (type == 3) {
                    err = ST_BLOCK_TYPE;
                    state = LS_DONE;
                    continue;
                }
                if (type == 1) {
                    // fixed code: sorted lists in closed form
                    auto fixed_sym = [](uint32_t e) -> uint32_t {
                        return e < 24 ? 256 + e : e < 168 ? e - 24 : e < 176 ? 280 + e - 168 : 144 + e - 176;
                    };
                    for (uint32_t w = 0; w < 72; ++w) {
                        uint32_t v = 0;
                        for (uint32_t j = 0; j < 4; ++j) v |= (fixed_sym(4 * w + j) & 255u) << (8 * j);
                        lw[w * LANES] = v;
                    }
                    for (uint32_t w = 0; w < 9; ++w) {
                        uint32_t v = 0;
                        for (uint32_t j = 0; j < 32; ++j) v |= (uint32_t)(fixed_sym(32 * w + j) >= 256) << j;
                        hw[w * LANES] = v;
                    }
                    for (uint32_t w = 0; w < 8; ++w) dw[w * LANES] = (4 * w) | (4 * w + 1) << 8 | (4 * w + 2) << 16 | (4 * w + 3) << 24;
                    (void)lane_code(lc, 288, [](int k) -> uint32_t { return k == 7 ? 24 : k == 8 ? 152 : k == 9 ? 112 : 0; });
                    (void)lane_code(dc, 32, [](int k) -> uint32_t { return k == 5 ? 32 : 0; });
                    state = LS_CODES;
                    continue;
                }
                // dynamic header up to the code-length code; the lengths are
                // read in the decode steps (LS_LENS), in parallel with the
                // other lanes' symbols
                nlen = lb_bits(b, 5, err) + 257;
                ndist = lb_bits(b, 5, err) + 1;
                const uint32_t ncode = lb_bits(b, 4, err) + 4;
                if (err == ST_OK && (nlen > 286 || ndist > 30)) err = ST_BAD_COUNTS;
                uint64_t cl = 0; // code-length code lengths, 3 bits per symbol 0..18
#pragma unroll
                for (uint32_t i = 0; i < 19; ++i) {
                    if (err == ST_OK && i < ncode) {
                        LB_REFILL_SYNC(b);
                        cl |= (uint64_t)lb_bits(b, 3, err) << (3 * clen_order(i));
                    }
                }
                if (err == ST_OK) {
                    // the code goes in dc (its limits flatten past length 7),
                    // its symbols sorted by (length, symbol) in the distance
                    // list's bytes
                    F8 c8 = {0, 0};
#pragma unroll
                    for (uint32_t s = 0; s < 19; ++s) f8_add(c8, (uint32_t)(cl >> (3 * s)) & 7, 1);
                    const int e = lane_code(dc, 19, [&](int k) -> uint32_t { return k < 8 ? f8_get(c8, (uint32_t)k) : 0; });
                    if (e != 0) err = ST_BAD_TABLE;
                    F8 o8 = {0, 0};
                    uint32_t run = 0;
#pragma unroll
                    for (uint32_t k = 1; k < 8; ++k) {
                        f8_add(o8, k, run);
                        run += f8_get(c8, k);
                    }
#pragma unroll
                    for (uint32_t s = 0; s < 19; ++s) {
                        const uint32_t L = (uint32_t)(cl >> (3 * s)) & 7;
                        if (L) {
                            const uint32_t p = f8_get(o8, L);
                            f8_add(o8, L, 1);
                            LANE_BYTE(dw, p < 32 ? p : 31) = (uint8_t)s;
                        }
                    }
                }
                if (err != ST_OK) {
                    state = LS_DONE;
                    continue;
                }
                total = nlen + ndist;
                index = 0;
                prev = 0;
                len256 = 0;
                pass = 0;
                cll = F16{0, 0, 0, 0};
                dn0 = dn1 = dn2 = dn3 = 0;
                lbit = in_len * 8 - b.avail;
                state = LS_LENS;
        } else if (state == LS_TABLES) {
            // the end of a pass over the code lengths
            if (pass == 0) {
                // puff's table checks, in its order
                if (len256 == 0) err = ST_BAD_TABLE;
                uint32_t sll = 0;
#pragma unroll
                for (uint32_t k = 1; k < 16; ++k) sll += f16_get(cll, k);
                const int e = lane_code(lc, nlen, [&](int k) -> uint32_t { return f16_get(cll, (uint32_t)k); });
                if (err == ST_OK && e != 0 && (e < 0 || sll != 1)) err = ST_BAD_TABLE;
                const F8 cd = dist_counts(dn0, dn1, dn2, dn3, ndist);
                uint32_t sd = 0;
                int dleft = 1;
#pragma unroll
                for (uint32_t k = 1; k < 16; ++k) {
                    sd += f8_get(cd, k);
                    dleft = (dleft << 1) - (int)f8_get(cd, k);
                }
                if (sd == 0) dleft = 0;
                if (err == ST_OK && dleft != 0 && (dleft < 0 || sd != 1)) err = ST_BAD_TABLE;
                if (err != ST_OK) {
                    state = LS_DONE;
                    continue;
                }
                // counts -> first list entry per code length, for pass 1
                F16 o = {0, 0, 0, 0};
                uint32_t r = 0;
#pragma unroll
                for (uint32_t k = 1; k < 16; ++k) {
                    f16_add(o, k, r);
                    r += f16_get(cll, k);
                }
                cll = o;
                for (uint32_t w = 0; w < 9; ++w) hw[w * LANES] = 0;
                lb_seek(lbit);
                index = 0;
                prev = 0;
                pass = 1;
                state = LS_LENS;
            } else {
                // the distance list from the lengths in dn0..dn3 (the
                // code-length symbols in dw are no longer needed)
                const F8 cd = dist_counts(dn0, dn1, dn2, dn3, ndist);
                F8 od = {0, 0};
                uint32_t r = 0;
#pragma unroll
                for (uint32_t k = 1; k < 16; ++k) {
                    f8_add(od, k, r);
                    r += f8_get(cd, k);
                }
#pragma unroll
                for (uint32_t j = 0; j < 30; ++j) {
                    const uint32_t dn = j < 8 ? dn0 : j < 16 ? dn1 : j < 24 ? dn2 : dn3;
                    const uint32_t L = (dn >> (4 * (j & 7))) & 15;
                    if (j < ndist && L) {
                        const uint32_t q = f8_get(od, L);
                        f8_add(od, L, 1);
                        LANE_BYTE(dw, q < 32 ? q : 31) = (uint8_t)j;
                    }
                }
                (void)lane_code(dc, ndist, [&](int k) -> uint32_t { return f8_get(cd, (uint32_t)k); });
                state = LS_CODES;
            }
        }
#pragma unroll
        for (uint32_t sub = 0; sub < LSTEPS; ++sub) {
            LB_REFILL_SEL(b); // harmless for a lane at a header
            // a symbol needs up to 20 bits, then (a match) a refill and 28
            // (a second literal 15); a lane without them waits for the next
            // pass
            if ((state != LS_CODES && state != LS_LENS) || b.bc < 32 || (b.bc < 48 && b.k == 8)) continue;
            // One symbol, and its distance when it is a length or a second
            // literal after a literal, without
            // branches: every lane runs the same instructions, flags pick
            // what each result means. Only the stores and errors branch. A
            // lane reading code lengths uses only the second code (dc holds
            // the code-length code, dw its symbols).
            const bool lens = state == LS_LENS;
            const uint32_t avail0 = b.avail;
            const uint32_t lo0 = (uint32_t)b.bb;
            uint32_t L1;
            const uint32_t idx1 = lane_find(lc, __builtin_bitreverse32(lo0) >> 17, L1);
            const bool ok1 = !lens && L1 != 0 && L1 <= avail0;
            const uint32_t e1 = idx1 < 288 ? idx1 : 287;
            const uint32_t sym = LANE_BYTE(lw, e1) | (((hw[(e1 >> 5) * LANES] >> (e1 & 31)) & 1u) << 8);
            const uint32_t n1 = ok1 ? L1 : 0;
            const bool lit = ok1 && sym < 256 && pos < cap;
            const bool eob = ok1 && sym == 256;
            const bool lenc = ok1 && sym - 257 < 29;
            const uint32_t lcode = lenc ? sym - 257 : 0;
            const uint32_t xb = len_extra(lcode);
            const bool okx = lenc && n1 + xb <= avail0;
            const uint32_t mlen = len_base(lcode) + __builtin_amdgcn_ubfe(lo0, n1, xb);
            const uint32_t c1 = n1 + (okx ? xb : 0);
            b.bb >>= c1;
            b.bc -= c1;
            b.avail = avail0 - c1;
            LB_REFILL_SEL(b);
            const uint32_t avail2 = b.avail;
            const uint32_t lo2 = (uint32_t)b.bb;
            // The second code: after a length its distance, in LS_LENS a
            // code-length symbol (both from dc and dw), after a literal the
            // next literal/length code (lc, lw), taken only when it is a
            // literal whose byte shares the first literal's output word (one
            // word closes per step at most)
            const bool tb = lit;
            LaneCode c2c;
#pragma unroll
            for (uint32_t i = 0; i < LQ; ++i) c2c.q[i] = tb ? lc.q[i] : dc.q[i];
            c2c.lim15 = tb ? lc.lim15 : dc.lim15;
            uint32_t L2;
            const uint32_t idx2 = lane_find(c2c, __builtin_bitreverse32(lo2) >> 17, L2);
            const bool ok2 = (okx || lens) && L2 != 0 && L2 <= avail2;
            const uint32_t e2 = min(idx2, tb ? 287u : 31u);
            const uint32_t ds = LANE_BYTE(tb ? lw : dw, e2);
            const bool lit2 = tb && L2 != 0 && L2 <= avail2 && ((hw[(e2 >> 5) * LANES] >> (e2 & 31)) & 1u) == 0 &&
                              pos + 1 < cap && ((pos + h + 1) & 3) != 0;
            // a distance code, or a code-length symbol (0..18): its extra
            // bits and base give the distance or the repeat count
            const bool okd = ok2 && (lens || ds < 30);
            const uint32_t dcode = okd && !lens ? ds : 0;
            // repeat codes 16, 17, 18: 2, 3, 7 extra bits on 3, 3, 11; the
            // literal lengths 0-15 none on 1 (nibble tables, then a bitwise
            // select on `lens`, which the compiler keeps branch-free)
            const uint32_t t16 = ds - 16;
            const uint32_t xl = t16 < 3 ? (0x732u >> (4 * t16)) & 15u : 0u;
            const uint32_t bl = t16 < 3 ? (0xB33u >> (4 * t16)) & 15u : 1u;
            const uint32_t lm = 0u - (uint32_t)lens;
            const uint32_t xd = (xl & lm) | (dist_extra(dcode) & ~lm);
            const uint32_t vbase = (bl & lm) | (dist_base(dcode) & ~lm);
            const bool rep0 = lens && ds == 16 && index == 0; // checked before the extra bits
            const bool okxd = okd && !rep0 && L2 + xd <= avail2;
            const uint32_t val = vbase + __builtin_amdgcn_ubfe(lo2, L2, xd);
            const bool match = !lens && okxd && val <= pos && pos + mlen <= cap;
            const bool clok = lens && okxd && index + val <= total;
            const uint32_t c2 = match || clok ? L2 + xd : lit2 ? L2 : 0;
            b.bb >>= c2;
            b.bc -= c2;
            b.avail = avail2 - c2;
            if (!(lit || eob || match || clok)) {
                // the first check that failed, in puff's order
                if (lens) {
                    if (!ok2)
                        err = L2 == 0 && avail2 >= 15 ? ST_BAD_CODE : ST_TRUNCATED;
                    else if (rep0)
                        err = ST_BAD_LENGTH_REPEAT;
                    else if (!okxd)
                        err = ST_TRUNCATED;
                    else
                        err = ST_BAD_LENGTH_REPEAT;
                } else if (!ok1) {
                    err = L1 == 0 && avail0 >= 15 ? ST_BAD_CODE : ST_TRUNCATED;
                } else if (sym < 256) {
                    err = ST_OUTPUT_OVERFLOW;
                } else if (!lenc) {
                    err = ST_BAD_CODE;
                } else if (!okx) {
                    err = ST_TRUNCATED;
                } else if (!ok2) {
                    err = L2 == 0 && avail2 >= 15 ? ST_BAD_DISTANCE : ST_TRUNCATED;
                } else if (!okd) {
                    err = ST_BAD_DISTANCE;
                } else if (!okxd) {
                    err = ST_TRUNCATED;
                } else if (val > pos) {
                    err = ST_BAD_DISTANCE;
                } else {
                    err = ST_OUTPUT_OVERFLOW;
                }
                state = LS_DONE;
            }
            {
                // a literal in a new window closes the open one
                const uint32_t rel = pos + h;
                const bool close = lit && (rel >> 2) != wcur;
                ls_i[sub] = close ? wcur : NONE;
                ls_v[sub] = wv;
                wv = close ? 0u : wv;
                wcur = lit ? rel >> 2 : wcur;
                wv |= lit ? sym << (8 * (rel & 3)) : 0u;
                wv |= lit2 ? ds << (8 * ((rel + 1) & 3)) : 0u;
            }
            if (match) {
                uint32_t gap = pos - mark;
                while (gap >= REC_SKIP) {
                    rec[nrec++] = REC_SKIP << 23;
                    gap -= REC_SKIP;
                }
                rs_i[sub] = nrec++;
                rs_v[sub] = (gap << 23) | ((mlen - 3) << 15) | (val - 1);
            }
            pos += lit ? 1 + (uint32_t)lit2 : match ? mlen : 0;
            mark = match ? pos : mark;
            if (eob) state = last ? LS_DONE : LS_HEADER;
            if (clok) {
                // val code lengths of clen; the first nll are literal/length
                // symbols, the rest distance symbols
                const uint32_t clen = ds < 16 ? ds : ds == 16 ? prev : 0;
                const uint32_t nll = index < nlen ? (val < nlen - index ? val : nlen - index) : 0;
                if (clen) {
                    const uint32_t q = f16_get(cll, clen);
                    f16_add(cll, clen, nll); // pass 0: counts; pass 1: next entries
                    if (pass) {
                        // symbols of one length take consecutive list entries
                        for (uint32_t k = 0; k < nll; ++k) {
                            const uint32_t e = q + k < 288 ? q + k : 287;
                            const uint32_t sy = index + k;
                            LANE_BYTE(lw, e) = (uint8_t)sy;
                            if (sy >= 256) atomicOr(&hw[(e >> 5) * LANES], 1u << (e & 31));
                        }
                    } else {
                        for (uint32_t k = nll; k < val; ++k) {
                            const uint32_t j = index + k - nlen;
                            const uint32_t v = clen << (4 * (j & 7));
                            dn0 |= j < 8 ? v : 0u;
                            dn1 |= j - 8 < 8 ? v : 0u;
                            dn2 |= j - 16 < 8 ? v : 0u;
                            dn3 |= j - 24 < 8 ? v : 0u;
                        }
                    }
                }
                if (index <= 256 && 256 < index + val) len256 = clen;
                prev = clen;
                index += val;
                if (index == total) state = LS_TABLES;
            }
        }
    }
    flush_pass();
    if (wcur != NONE) put_window(wcur, wv, NONE);
    a.status[m] = err;
    a.produced[m] = pos;
    a.crc[m] = 0;
    a.nrec[m] = nrec;
#undef LANE_BYTE
}

// Four bytes of a partial block, skipping those at or past `end`.
__device__ __forceinline__ uint32_t part4(const FR_GLOBAL uint8_t *p, const FR_GLOBAL uint8_t *end)
{
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) v |= p + j < end ? (uint32_t)p[j] << (8 * j) : 0u;
    return v;
}

// Output window of a member for one round: lane l holds the 16-byte aligned
// block at ((g + from) & ~15) + 16 l, so the wave covers at least positions
// [from, from + WIN - 15). Blocks past position P are not read; the block
// that straddles P is read byte by byte.
constexpr uint32_t WIN = 1024;
__device__ __forceinline__ uint4 load_window(const FR_GLOBAL uint8_t *g, uint32_t from, uint32_t P, uint32_t lane)
{
    const FR_GLOBAL uint8_t *blk = g + from - ((uintptr_t)(g + from) & 15) + 16 * lane;
    const FR_GLOBAL uint8_t *end = g + P;
    uint4 v = {0, 0, 0, 0};
    if (blk + 16 <= end) {
        v = gload16(blk);
    } else if (blk < end) {
        v.x = part4(blk, end);
        v.y = part4(blk + 4, end);
        v.z = part4(blk + 8, end);
        v.w = part4(blk + 12, end);
    }
    return v;
}

// flush() on a copy of the output state, returned by value: callers that
// would otherwise pass their state by reference to an out-of-line call keep it
// in registers.
template <uint32_t RING>
__device__ __attribute__((noinline)) Out flush_copy(const uint8_t *ring, const uint32_t *crctab, Out o, uint32_t lane,
                                                    uint32_t lane_factor, const uint32_t *x2n)
{
    flush<RING>(ring, crctab, o, lane, lane_factor, x2n);
    return o;
}

// Inclusive prefix sum over the wave with DPP row shifts and row broadcasts
// (no LDS round trips).
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t x)
{
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false); // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false); // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false); // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false); // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false); // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false); // row_bcast:31
    return x;
}

// Write-back of exactly FAST_CHUNK bytes from a 16-byte aligned output
// position: lane l moves the 32 bytes at 32 l to HBM and takes their CRC with
// the slice-by-8 tables (t8[k][b] = CRC of byte b followed by k zero bytes;
// each 8-byte step takes the CRC so far into its first four bytes). Each lane
// folds its pieces across write-backs by Horner's rule, acc = acc * X ^ crc
// with X = x^(8 FAST_CHUNK), the product by X read from tables (tx[k][b] =
// (b at byte k) * X); the lanes' sums fold with x^(8*32*(63-l)) once, at the
// member's end.
constexpr uint32_t FAST_CHUNK = 2048;
__device__ __forceinline__ uint32_t slice8(const uint32_t *t8, uint32_t r, uint32_t lo, uint32_t hi)
{
    lo ^= r;
    return t8[7 * 256 + (lo & 0xFF)] ^ t8[6 * 256 + ((lo >> 8) & 0xFF)] ^ t8[5 * 256 + ((lo >> 16) & 0xFF)] ^
           t8[4 * 256 + (lo >> 24)] ^ t8[3 * 256 + (hi & 0xFF)] ^ t8[2 * 256 + ((hi >> 8) & 0xFF)] ^
           t8[1 * 256 + ((hi >> 16) & 0xFF)] ^ t8[hi >> 24];
}
template <uint32_t RING>
__device__ __forceinline__ void flush_fast(const uint8_t *ring, const uint32_t *t8, const uint32_t *tx, Out &o,
                                           uint32_t lane, uint32_t &acc, uint32_t &nfast)
{
    using G = Geometry<RING>;
    WFENCE();
    const uint32_t p = o.flushed + 32 * lane;
    const uint4 v0 = *reinterpret_cast<const uint4 *>(ring + ((o.skew + p) & G::RMASK));
    const uint4 v1 = *reinterpret_cast<const uint4 *>(ring + ((o.skew + p + 16) & G::RMASK));
    gstore16(o.g + p, v0);
    gstore16(o.g + p + 16, v1);
    uint32_t r = slice8(t8, 0, v0.x, v0.y);
    r = slice8(t8, r, v0.z, v0.w);
    r = slice8(t8, r, v1.x, v1.y);
    r = slice8(t8, r, v1.z, v1.w);
    acc = tx[acc & 0xFF] ^ tx[256 + ((acc >> 8) & 0xFF)] ^ tx[512 + ((acc >> 16) & 0xFF)] ^ tx[768 + (acc >> 24)] ^ r;
    ++nfast;
    o.flushed += FAST_CHUNK;
}

// Phase 2: records of one member, 64 at a time, into the ring. Each round's
// records and output window are fetched one round ahead, into one of two
// register sets used in turn (so a fetched value is never copied, which would
// wait for it).
template <uint32_t RING, uint32_t NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void fr_inflate_resolve_kernel(InflateArgs a)
{
    using G = Geometry<RING>;
    // a round covers at most this many output bytes (one record spans <= 768)
    constexpr uint32_t SPAN = WIN - 16;
    static_assert(SPAN >= REC_SKIP + 258, "window too small for a record");
    static_assert(SPAN + WIN <= RING / 2, "window writes must stay clear of unflushed bytes");
    static_assert(FAST_CHUNK + SPAN + WIN <= RING, "fast write-back must stay clear of the window writes");
    static_assert(SPAN <= 64 * G::SEG && FAST_CHUNK <= 64 * G::SEG,
                  "the last write-back (under FAST_CHUNK bytes after a fast one) must fit flush()");
    __shared__ uint32_t t8[8 * 256]; // t8[0] is the byte-at-a-time table
    __shared__ uint32_t tx[4 * 256]; // products by x^(8 FAST_CHUNK), per register byte
    __shared__ __attribute__((aligned(16))) uint8_t rings[NW][RING];
    __shared__ uint4 stage_all[NW][64][2]; // per lane: the 32 aligned bytes around a far match source
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < 256; i += 64 * NW) {
        uint32_t c = i;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        t8[i] = c;
    }
    const uint32_t x_chunk = fr_xpow8(a.x2n, FAST_CHUNK);
    for (uint32_t i = threadIdx.x; i < 4 * 256; i += 64 * NW) tx[i] = fr_mulmod((i & 255u) << (8 * (i >> 8)), x_chunk);
    __syncthreads();
    for (uint32_t k = 1; k < 8; ++k) {
        for (uint32_t i = threadIdx.x; i < 256; i += 64 * NW) {
            const uint32_t c = t8[(k - 1) * 256 + i];
            t8[k * 256 + i] = (c >> 8) ^ t8[c & 0xFF];
        }
        __syncthreads();
    }
    const uint32_t *crctab = t8;
    uint8_t *ring = rings[wave];
    const uint32_t lane_factor = fr_xpow8(a.x2n, (uint64_t)G::SEG * (63 - lane));
    const uint32_t lane_factor32 = fr_xpow8(a.x2n, (uint64_t)32 * (63 - lane));
    uint32_t iter = 0;
    for (;;) {
        uint32_t m;
        if (a.queue) {
            uint32_t q = 0;
            if (lane == 0) q = atomicAdd(a.queue, 1u);
            m = uni(q);
        } else {
            m = (blockIdx.x + iter * gridDim.x) * NW + wave;
        }
        ++iter;
        if (m >= a.n) break;
        if (a.order) m = uni(a.order[m]);
        if (uni(a.status[m]) != ST_OK) continue;
        const uint32_t P = uni(a.produced[m]);
        const uint32_t nr = uni(a.nrec[m]);
        const uint64_t out_off = uni64(a.out_off[m]);
        const uint32_t *rec = a.rec + rec_base(out_off, m);
        Out o;
        o.g = (FR_GLOBAL uint8_t *)(a.out + out_off);
        o.pos = 0;
        o.flushed = 0;
        o.visible = 0;
        o.cap = P;
        o.raw = 0;
        o.skew = (uint32_t)((uintptr_t)o.g & 16383);
        uint32_t ri = 0;
        bool bad = false;
        uint32_t acc = 0, nfast = 0; // this lane's fast write-back CRCs, and their count
        // One round: records `rc` (from ri) and window `wc` (from o.pos) are
        // in hand; those of the next round go to rn, wn.
        auto round = [&](const uint32_t rc, const uint4 wc, uint32_t &rn, uint4 &wn) __attribute__((always_inline)) {
            ri = uni(ri);
            o.pos = uni(o.pos);
            o.flushed = uni(o.flushed);
            // write-back: the bytes up to the first 16-byte aligned output
            // address, then aligned FAST_CHUNK pieces (fewer unflushed bytes
            // than flush() can carry remain at the member's end)
            if (((uintptr_t)(o.g + o.flushed) & 15) != 0) {
                if (o.pos - o.flushed >= 16) {
                    const uint32_t keep = o.pos;
                    o.pos = o.flushed + 16 - (uint32_t)((uintptr_t)(o.g + o.flushed) & 15);
                    o = flush_copy<RING>(ring, crctab, o, lane, lane_factor, a.x2n);
                    o.pos = keep;
                }
            } else if (o.pos - o.flushed >= FAST_CHUNK) {
                flush_fast<RING>(ring, t8, tx, o, lane, acc, nfast);
            }
            const uint32_t R0 = o.pos;
            const bool live = ri + lane < nr;
            const uint32_t gap = rc >> 23;
            const uint32_t len = !live || gap == REC_SKIP ? 0 : ((rc >> 15) & 255) + 3;
            const uint32_t dist = (rc & 0x7FFFu) + 1;
            const uint32_t inc = wave_incl_sum(live ? gap + len : 0);
            const uint32_t k = (uint32_t)__popcll(__ballot(live && inc <= SPAN));
            uint32_t R1;
            if (k == 0) {
                R1 = P - R0 < SPAN ? P : R0 + SPAN; // literals after the last record
            } else {
                R1 = R0 + (uint32_t)__builtin_amdgcn_readlane((int)inc, (int)k - 1);
            }
            if ((k == 0 && ri < nr) || R1 > P) {
                // records that do not fit the member (phase 1 never writes such)
                bad = true;
                return;
            }
            // the next round's records and window
            rn = ri + k + lane < nr ? rec[ri + k + lane] : 0;
            wn = load_window(o.g, R1, P, lane);
            // This round's matches. A source wholly in HBM is read as the one
            // or two aligned 16-byte blocks holding it (two requests, not one
            // per byte), issued here and used after the window and the ring
            // copies, which hide their latency.
            // ring slots are good from R0 + WIN - RING on: the window writes
            // of this and earlier rounds reach at most WIN bytes past R0
            const uint32_t lim = R0 + WIN > RING ? R0 + WIN - RING : 0;
            const bool match = lane < k && len > 0;
            const uint32_t dst = R0 + inc - len;
            const uint32_t src = dst - dist;
            const bool deferred = match && (src + len > R0 || len > TOK_SHORT);
            const bool far = match && !deferred && src + len <= lim;
            const bool near = match && !deferred && src >= lim;
            const uint32_t fhead = (uint32_t)((uintptr_t)(o.g + src) & 15);
            // this wave's write-backs, before it reads them from HBM
            if (__ballot(match && src < lim) != 0 && o.visible < o.flushed) {
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
                o.visible = o.flushed;
            }
            // left unset where not loaded: a zero fill of the registers would
            // have to wait for the loads of other lanes into them
            uint4 f0, f1;
            if (far) {
                const FR_GLOBAL uint8_t *blk = o.g + src - fhead;
                f0 = gload16(blk);
                if (fhead + len > 16) f1 = gload16(blk + 16);
            }
            // this round's window into the ring: whole blocks, but the first
            // block only from R0 on (the bytes before R0 are this member's
            // finished output, possibly not yet written back)
            {
                const uint32_t head = (uint32_t)((uintptr_t)(o.g + R0) & 15);
                const uint32_t S = R0 - head + 16 * lane;
                uint4 *slot = reinterpret_cast<uint4 *>(ring + ((o.skew + S) & G::RMASK));
                uint4 v = wc;
                if (lane == 0 && head) {
                    // keep the ring's bytes below R0: merge by byte masks
                    const uint4 old = *slot;
                    auto keep = [&](uint32_t word) -> uint32_t {
                        const int32_t n = (int32_t)head - (int32_t)(4 * word); // bytes of this word to keep
                        return n >= 4 ? ~0u : n <= 0 ? 0u : (1u << (8 * n)) - 1;
                    };
                    const uint32_t k0 = keep(0), k1 = keep(1), k2 = keep(2), k3 = keep(3);
                    v.x = (old.x & k0) | (wc.x & ~k0);
                    v.y = (old.y & k1) | (wc.y & ~k1);
                    v.z = (old.z & k2) | (wc.z & ~k2);
                    v.w = (old.w & k3) | (wc.w & ~k3);
                }
                *slot = v;
            }
            WFENCE();
            // Bytes from `B` at q (masked by M) to the ring at d: up to an
            // aligned ring word, whole words (two aligned source words
            // funnel-shifted), then the last bytes. Only bytes of this
            // match's span are written.
            auto copy = [&](const uint8_t *B, uint32_t M, uint32_t q) __attribute__((always_inline)) {
                const uint32_t d = o.skew + dst;
                const uint32_t pre = min((0u - d) & 3u, len);
                uint32_t i = 0;
                for (; i < pre; ++i) ring[(d + i) & G::RMASK] = B[(q + i) & M];
                for (; i + 4 <= len; i += 4) {
                    const uint32_t sp = q + i, al = sp & ~3u;
                    const uint32_t w0 = *reinterpret_cast<const uint32_t *>(B + (al & M));
                    const uint32_t w1 = *reinterpret_cast<const uint32_t *>(B + ((al + 4) & M));
                    *reinterpret_cast<uint32_t *>(ring + ((d + i) & G::RMASK)) =
                        __builtin_amdgcn_alignbit(w1, w0, (sp & 3) * 8);
                }
                for (; i < len; ++i) ring[(d + i) & G::RMASK] = B[(q + i) & M];
            };
            // sources in the ring lie before R0, which no write of this round
            // changes
            if (near) copy(ring, G::RMASK, o.skew + src);
            // sources in HBM, through this lane's 32-byte stage
            uint4 *stg = stage_all[wave][lane];
            if (far) {
                stg[0] = f0;
                stg[1] = f1;
                copy(reinterpret_cast<const uint8_t *>(stg), ~0u, fhead);
            }
            if (match && !deferred && !far && !near) {
                // straddles lim: byte by byte from HBM or the ring
                for (uint32_t i = 0; i < len; ++i) {
                    const uint32_t p = src + i;
                    ring[(o.skew + dst + i) & G::RMASK] = p >= lim ? ring[(o.skew + p) & G::RMASK] : o.g[p];
                }
            }
            WFENCE();
            uint64_t pending = __ballot(deferred);
            while (pending) {
                const uint32_t j = (uint32_t)__ffsll((unsigned long long)pending) - 1;
                pending &= pending - 1;
                const uint32_t d0 = (uint32_t)__builtin_amdgcn_readlane((int)dst, (int)j);
                const uint32_t L = (uint32_t)__builtin_amdgcn_readlane((int)len, (int)j);
                const uint32_t D = (uint32_t)__builtin_amdgcn_readlane((int)dist, (int)j);
                if (D >= L) {
                    for (uint32_t off = lane; off < L; off += 64) {
                        const uint32_t p = d0 - D + off;
                        const uint8_t v = p >= lim ? ring[(o.skew + p) & G::RMASK] : o.g[p];
                        ring[(o.skew + d0 + off) & G::RMASK] = v;
                    }
                } else {
                    uint32_t rr = lane % D;
                    const uint32_t step = 64u % D;
                    for (uint32_t off = lane; off < L; off += 64) {
                        ring[(o.skew + d0 + off) & G::RMASK] = ring[(o.skew + d0 - D + rr) & G::RMASK];
                        rr += step;
                        if (rr >= D) rr -= D;
                    }
                }
                WFENCE();
            }
            o.pos = R1;
            ri += k;
        };
        uint32_t rA = lane < nr ? rec[lane] : 0, rB = 0;
        uint4 wA = load_window(o.g, 0, P, lane), wB = {0, 0, 0, 0};
        for (;;) {
            if (!(ri < nr || o.pos < P) || bad) break;
            round(rA, wA, rB, wB);
            if (!(ri < nr || o.pos < P) || bad) break;
            round(rB, wB, rA, wA);
        }
        if (bad) {
            if (lane == 0) a.status[m] = ST_OUTPUT_OVERFLOW;
            continue;
        }
        if (((uintptr_t)(o.g + o.flushed) & 15) == 0 && o.pos - o.flushed >= FAST_CHUNK)
            flush_fast<RING>(ring, t8, tx, o, lane, acc, nfast);
        if (nfast) {
            // the fast write-backs' CRC after the bytes before them
            uint32_t r = fr_mulmod(acc, lane_factor32);
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) r ^= __shfl_xor(r, off, 64);
            o.raw = fr_mulmod(o.raw, fr_xpow8(a.x2n, (uint64_t)FAST_CHUNK * nfast)) ^ uni(r);
        }
        o = flush_copy<RING>(ring, crctab, o, lane, lane_factor, a.x2n);
        if (lane == 0) a.crc[m] = o.raw ^ fr_mulmod(0xFFFFFFFFu, fr_xpow8(a.x2n, o.pos)) ^ 0xFFFFFFFFu;
    }
}

// CRC32 of independent byte ranges data[off[m], off[m] + len[m]), one wave per
// range, four waves per workgroup. A range is a head of up to 15 bytes, a body
// of 16-byte aligned chunks and a tail of up to 15 bytes. Lane l takes body
// chunks l, l + 64, ...: each step the wave loads 1 KiB with one coalesced
// 16-byte load per lane; a chunk's raw CRC is sixteen slice-by-16 table
// lookups, and the lane folds its chunks by Horner's rule, multiplying its
// accumulator by x^(8*1024) (the 1 KiB between a lane's chunks) with four
// 256-entry tables. The lanes' accumulators, weighted by the bytes after their
// last chunk, and the head and tail CRCs fold into the range's CRC.
constexpr uint32_t CRC_WAVES = 4;

__global__ __launch_bounds__(64 * CRC_WAVES) void fr_crc32_kernel(const uint8_t *data, const uint64_t *off,
                                                                  const uint32_t *len, uint32_t n, uint32_t *crc,
                                                                  const uint32_t *x2n)
{
    __shared__ uint32_t s16[16][256]; // raw CRC of byte v followed by k zero bytes
    __shared__ uint32_t zt[4][256];   // (v << 8j) * x^(8*1024) mod P
    const uint32_t t = threadIdx.x;
    {
        uint32_t c = t;
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (FR_CRC_POLY & (0u - (c & 1u)));
        s16[0][t] = c;
    }
    __syncthreads();
    for (int k = 1; k < 16; ++k) {
        const uint32_t v = s16[k - 1][t];
        s16[k][t] = s16[0][v & 0xFF] ^ (v >> 8);
        __syncthreads();
    }
    const uint32_t X = fr_xpow8(x2n, 1024);
#pragma unroll
    for (int j = 0; j < 4; ++j) zt[j][t] = fr_mulmod(t << (8 * j), X);
    __syncthreads();
    const uint32_t lane = t & 63, w = t >> 6;
    for (uint32_t m = blockIdx.x * CRC_WAVES + w; m < n; m += gridDim.x * CRC_WAVES) {
        const uint64_t a = off[m];
        const uint32_t L = len[m];
        const uint64_t e = a + L;
        uint64_t A0 = (a + 15) & ~15ull, E0 = e & ~15ull;
        if (A0 > e) A0 = e;
        if (E0 < A0) E0 = A0;
        uint32_t part = 0;
        if (lane == 0) {
            uint32_t r = 0;
            for (uint64_t p = a; p < A0; ++p) r = s16[0][(r ^ data[p]) & 0xFF] ^ (r >> 8);
            part = fr_mulmod(r, fr_xpow8(x2n, e - A0));
        } else if (lane == 1) {
            uint32_t r = 0;
            for (uint64_t p = E0; p < e; ++p) r = s16[0][(r ^ data[p]) & 0xFF] ^ (r >> 8);
            part = r;
        }
        const uint64_t nc = (E0 - A0) / 16;
        const uint4 *body = reinterpret_cast<const uint4 *>(data + A0);
        uint32_t acc = 0;
        uint64_t last = ~0ull;
        for (uint64_t c = lane; c < nc; c += 64) {
            const uint4 v = body[c];
            const uint32_t q[4] = {v.x, v.y, v.z, v.w};
            uint32_t r = 0;
#pragma unroll
            for (int i = 0; i < 16; ++i) r ^= s16[15 - i][(q[i >> 2] >> (8 * (i & 3))) & 0xFF];
            acc = zt[0][acc & 0xFF] ^ zt[1][(acc >> 8) & 0xFF] ^ zt[2][(acc >> 16) & 0xFF] ^ zt[3][acc >> 24] ^ r;
            last = c;
        }
        if (last != ~0ull) part ^= fr_mulmod(acc, fr_xpow8(x2n, 16 * (nc - 1 - last) + (e - E0)));
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) part ^= __shfl_xor(part, o, 64);
        if (lane == 0) crc[m] = part ^ fr_mulmod(0xFFFFFFFFu, fr_xpow8(x2n, L)) ^ 0xFFFFFFFFu;
    }
}

} // namespace

hipError_t launch_crc32_blocks(const uint8_t *data, const uint64_t *off, const uint32_t *len, uint32_t n,
                               uint32_t *crc, const uint32_t *x2n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint32_t grid = (n + CRC_WAVES - 1) / CRC_WAVES;
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(fr_crc32_kernel, dim3(grid), dim3(64 * CRC_WAVES), 0, s, data, off, len, n, crc, x2n);
    return hipGetLastError();
}

template <uint32_t RING, uint32_t NW, int TOK = 0>
static hipError_t launch_variant(const InflateArgs &a, uint32_t waves, hipStream_t stream)
{
    uint32_t grid = (waves + NW - 1) / NW;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL((fr_inflate_kernel<RING, NW, TOK>), dim3(grid), dim3(64 * NW), 0, stream, a);
    return hipGetLastError();
}

template <uint32_t RING, uint32_t NW>
static hipError_t launch_two_phase(const InflateArgs &a, uint32_t waves, hipStream_t stream)
{
    if (!a.rec || !a.nrec) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fr_inflate_lanes_kernel<1>, dim3((a.n + 63) / 64), dim3(64), 0, stream, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    uint32_t grid = (waves + NW - 1) / NW;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL((fr_inflate_resolve_kernel<RING, NW>), dim3(grid), dim3(64 * NW), 0, stream, a);
    return hipGetLastError();
}

uint64_t inflate_scratch_bytes(int schedule, uint64_t out_bytes, uint64_t n)
{
    if (schedule < 0 || schedule >= N_INFLATE_SCHEDULES || INFLATE_SCHEDULES[schedule].tokens != 3) return 0;
    const uint64_t head = (4 * n + 255) & ~255ull;
    return head + 4 * ((out_bytes * 11) / 32 + 2 * n + 4);
}

void inflate_scratch_bind(InflateArgs &a, void *scratch, uint64_t n)
{
    a.nrec = (uint32_t *)scratch;
    a.rec = scratch ? (uint32_t *)((uint8_t *)scratch + ((4 * n + 255) & ~255ull)) : nullptr;
}

const InflateSchedule INFLATE_SCHEDULES[] = {
    {32768, 1, 0}, {16384, 1, 0}, {16384, 2, 0}, {8192, 1, 0}, {8192, 2, 0}, {8192, 4, 0}, {4096, 2, 0}, {4096, 4, 0},
    {8192, 2, 1}, {8192, 4, 1}, {4096, 2, 1}, {4096, 4, 1},
    {8192, 2, 2}, {8192, 4, 2}, {4096, 2, 2}, {4096, 4, 2},
    {4096, 4, 3}, {4096, 2, 3}, {8192, 4, 3}, {8192, 2, 3},
};
const int N_INFLATE_SCHEDULES = sizeof(INFLATE_SCHEDULES) / sizeof(INFLATE_SCHEDULES[0]);

uint32_t inflate_lds_bytes(int schedule)
{
    const InflateSchedule &s = INFLATE_SCHEDULES[schedule];
    if (s.tokens == 3) return 16384 + s.waves_per_group * (s.ring + 2048); // the copy phase: CRC tables, rings, stages
    uint32_t per_wave = 0;
    switch (s.ring) {
    case 32768: per_wave = sizeof(WaveMem<32768>); break;
    case 16384: per_wave = sizeof(WaveMem<16384>); break;
    case 8192: per_wave = sizeof(WaveMem<8192>); break;
    default: per_wave = sizeof(WaveMem<4096>); break;
    }
    return 1024 + s.waves_per_group * per_wave;
}

// Resident waves per CU of a schedule: LDS-limited (160 KiB per CU), and at most
// 4 waves per SIMD for the kernel's register budget.
uint32_t inflate_waves_per_cu(int schedule)
{
    uint32_t groups = (160u * 1024u) / inflate_lds_bytes(schedule);
    uint32_t w = groups * INFLATE_SCHEDULES[schedule].waves_per_group;
    // register budget: 4 waves per SIMD
    return w > 16 ? 16 : (w ? w : 1);
}

hipError_t launch_inflate(const InflateArgs &a, uint32_t waves, hipStream_t stream, int schedule)
{
    if (a.n == 0) return hipSuccess;
    // The two-phase form decodes one member per lane: a batch too small to
    // fill the GPU that way takes a member's whole decode time, so it runs the
    // lane-speculative one-wave form with the same ring instead (same bytes).
    if (schedule >= 0 && schedule < N_INFLATE_SCHEDULES && INFLATE_SCHEDULES[schedule].tokens == 3 &&
        a.n < a.two_phase_min)
        schedule = INFLATE_SCHEDULES[schedule].ring == 4096 ? 15 : 13;
    switch (schedule) {
    case 0: return launch_variant<32768, 1>(a, waves, stream);
    case 1: return launch_variant<16384, 1>(a, waves, stream);
    case 2: return launch_variant<16384, 2>(a, waves, stream);
    case 3: return launch_variant<8192, 1>(a, waves, stream);
    case 4: return launch_variant<8192, 2>(a, waves, stream);
    case 5: return launch_variant<8192, 4>(a, waves, stream);
    case 6: return launch_variant<4096, 2>(a, waves, stream);
    case 7: return launch_variant<4096, 4>(a, waves, stream);
    case 8: return launch_variant<8192, 2, 1>(a, waves, stream);
    case 9: return launch_variant<8192, 4, 1>(a, waves, stream);
    case 10: return launch_variant<4096, 2, 1>(a, waves, stream);
    case 11: return launch_variant<4096, 4, 1>(a, waves, stream);
    case 12: return launch_variant<8192, 2, 2>(a, waves, stream);
    case 13: return launch_variant<8192, 4, 2>(a, waves, stream);
    case 14: return launch_variant<4096, 2, 2>(a, waves, stream);
    case 15: return launch_variant<4096, 4, 2>(a, waves, stream);
    case 16: return launch_two_phase<4096, 4>(a, waves, stream);
    case 17: return launch_two_phase<4096, 2>(a, waves, stream);
    case 18: return launch_two_phase<8192, 4>(a, waves, stream);
    case 19: return launch_two_phase<8192, 2>(a, waves, stream);
    default: return hipErrorInvalidValue;
    }
}

} // namespace frhip
