// deflate.hip -- BGZF blocks compressed on the device.
//
// The reference's device writer compresses with nvCOMP (a closed library) and
// frames its output with crates/fritillaria-cuda/kernels/bgzf_frame.cu. This is
// a DEFLATE compressor of its own (RFC 1951), one chunk per workgroup, with
// the chunk, its hash links and a hash table resident in LDS:
//   1. match    in rounds of 2,048 positions. Links: a hash chain over the
//               3-byte hashes, kept in a ring of the last 32 KiB + one round
//               of positions (prev[p % RING]: the previous position with the
//               same hash), linked in sub-rounds of 512: a sub-round's
//               positions of one hash are ranked from the highest down by
//               atomicMax (up to 8 of them linked in order) and the lowest
//               linked one continues to the last position of an earlier
//               sub-round, so the chains do not depend on thread order.
//               Pre-pass: each thread's 4 consecutive positions start from the
//               previous position's match (length - 1, same distance; final
//               when 32 or longer) and distances 1 to 4. Chain pass: lanes
//               take the round's other positions from a workgroup queue and
//               walk up to 128 links, stopping at a 128-byte match (zlib level
//               6's limits); a link is compared in full only when the 4 bytes
//               ending at the best length so far match, and every compare
//               reads the chunk in LDS. A 3-byte match farther than 4 KiB is
//               dropped, as zlib's TOO_FAR rule does. Lengths and distances
//               go to per-workgroup scratch.
//   2. parse    32 threads, one 2 KiB segment each, walk their segment with
//               zlib's lazy rule (a match shorter than 32 is deferred when the
//               next position has a longer one); matches stop at segment ends,
//               so the segments are independent.
//   3. codes    per dynamic block of 4096 symbols (statistics change along a
//               chunk: names, bases, qualities, kinetics), one wave per block,
//               8 blocks at a time: literal/length and distance histograms,
//               length-limited (15 bits) Huffman codes, the code-length code
//               (7 bits), the block header with runs of code lengths
//               (16/17/18), and the block's size;
//   4. emit     every symbol's bit offset from a workgroup prefix sum, its
//               bits OR-ed into an LDS staging area, block after block.
//   5. frame    crates/fritillaria-core/src/compress.rs:115-127 decides:
//               an empty chunk is the canonical empty block (so an empty
//               chunk frames to the EOF marker), a stream that would not beat
//               storing (or would not fit) is a stored block; the 18-byte
//               header with BC = block size - 1, the body, CRC32 and ISIZE go
//               to the chunk's 64 KiB slot.
// Output is not byte-identical to any other compressor's (no two DEFLATE
// compressors are); what holds is the round trip, the framing, and a ratio.
#include <hip/hip_runtime.h>

#include "deflate_internal.h"

namespace frhip {
namespace {

constexpr uint32_t DT = 512;
constexpr uint32_t NW = DT / 64;
constexpr uint32_t HBITS = 12, NBUCKET = 1u << HBITS;
constexpr uint32_t CHAIN = 128, NICE = 128; // zlib level 6's max_chain and nice_length
constexpr uint32_t KLINK = 8;               // same-hash positions of one round linked into the chain
constexpr uint32_t SEG = 2048, NSEG = 65536 / SEG;
constexpr uint32_t BSYM = 4096; // symbols per dynamic block
constexpr uint32_t WINDOW = 32768, MAX_LEN = 257, LAZY_MAX = 32, TOO_FAR = 4096;
constexpr uint16_t NO_POS = 0xFFFF;
constexpr uint32_t MATCH = 0x80000000u;
constexpr uint32_t MAX_CL = 288 + 32; // code-length symbols at most

// Code-length code order (RFC 1951 3.2.7).
__constant__ uint8_t CL_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

struct HuffWork {
    uint32_t wf[2 * 288];
    uint16_t parent[2 * 288];
    uint16_t order[288];
    uint8_t depth[2 * 288];
    uint32_t bl_count[16];
    uint32_t m;
};

// One dynamic block's codes and header.
struct Tables {
    uint32_t f_ll[288], f_d[32], f_cl[19];
    uint8_t l_ll[288], l_d[32], l_cl[19];
    uint16_t c_ll[288], c_d[32], c_cl[19];
    uint16_t cl_sym[MAX_CL]; // code-length symbol | extra << 5
    uint32_t n_cl, hlit, hdist, hclen, header_bits, sym_bits;
    HuffWork w;
};

// Positions per matching round (4 consecutive ones per thread), and the hash
// links kept while matching: the window behind a round, and the round.
constexpr uint32_t ROUND = 4 * DT;
constexpr uint32_t RING = WINDOW + ROUND;

struct Lds {
    union {
        struct {
            uint8_t data[65536 + 64]; // the chunk, zero past its end
            uint16_t prev[RING];      // the hash chain link of position p, at p % RING (NO_POS: none)
        } m;
        struct {
            uint8_t lens[65536]; // match length - 2 per position (0: none)
            uint32_t out[16384]; // the deflate stream being assembled
        } e;
    } b;
    union {
        struct {
            uint32_t stamp[NBUCKET]; // (round, iteration, offset) of the latest winner per hash
            uint16_t head[NBUCKET];  // last position + 1 per hash, of earlier rounds
        } ch;
    } c;
    uint32_t seg_count[NSEG], seg_base[NSEG + 1];
    uint64_t scan[NW];
    uint32_t chunk, n, framing, qnext;
    uint64_t total_bits;
};
static_assert(sizeof(Lds) <= 160 * 1024, "LDS budget of one workgroup");
// the code tables of NW blocks are built together in the match lengths' space
static_assert(NW * sizeof(Tables) <= 65536, "block tables must fit the lengths' space");

// Wave-local ordering of LDS accesses (a wave's lanes share one program order).
#define WSYNC()                                                                                                   \
    do {                                                                                                           \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                                                     \
        __builtin_amdgcn_wave_barrier();                                                                           \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                                                     \
    } while (0)

__device__ inline uint32_t hash3(const uint8_t *d)
{
    const uint32_t v = (uint32_t)d[0] | ((uint32_t)d[1] << 8) | ((uint32_t)d[2] << 16);
    return (v * 2654435761u) >> (32 - HBITS);
}

// Length of the common prefix of two 16-byte strings.
__device__ inline uint32_t prefix16(uint4 x, uint4 y)
{
    const uint32_t d0 = x.x ^ y.x, d1 = x.y ^ y.y, d2 = x.z ^ y.z, d3 = x.w ^ y.w;
    if (d0) return (uint32_t)__builtin_ctz(d0) >> 3;
    if (d1) return 4 + ((uint32_t)__builtin_ctz(d1) >> 3);
    if (d2) return 8 + ((uint32_t)__builtin_ctz(d2) >> 3);
    if (d3) return 12 + ((uint32_t)__builtin_ctz(d3) >> 3);
    return 16;
}

// 4 and 16 bytes at d + off in LDS, from aligned dwords (the chunk's copy
// is padded, so the dword after the last one wanted is always there).
__device__ inline uint32_t lds4(const uint8_t *d, uint32_t off)
{
    const uint32_t a = off & 3;
    const uint32_t *w = (const uint32_t *)(d + (off - a));
    return __builtin_amdgcn_alignbyte(w[1], w[0], a);
}
__device__ inline uint4 lds16(const uint8_t *d, uint32_t off)
{
    const uint32_t a = off & 3;
    const uint32_t *w = (const uint32_t *)(d + (off - a));
    const uint32_t w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3], w4 = w[4];
    uint4 v;
    v.x = __builtin_amdgcn_alignbyte(w1, w0, a);
    v.y = __builtin_amdgcn_alignbyte(w2, w1, a);
    v.z = __builtin_amdgcn_alignbyte(w3, w2, a);
    v.w = __builtin_amdgcn_alignbyte(w4, w3, a);
    return v;
}

// A match distance written earlier in the launch by another thread.
__device__ inline uint32_t ld_dist(const uint16_t *dist, uint32_t pos)
{
    const uint32_t w = __hip_atomic_load((const uint32_t *)dist + (pos >> 1), __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return (pos & 1) ? w >> 16 : w & 0xFFFF;
}

// Length code, extra bits and extra value of a match length 3..258.
__device__ inline void len_code(uint32_t L, uint32_t &code, uint32_t &nx, uint32_t &xv)
{
    const uint32_t x = L - 3;
    if (L == 258) {
        code = 285;
        nx = 0;
        xv = 0;
    } else if (x < 8) {
        code = 257 + x;
        nx = 0;
        xv = 0;
    } else {
        const uint32_t l = 31 - __clz(x); // 3..7
        const uint32_t sub = (x >> (l - 2)) & 3;
        code = 265 + 4 * (l - 3) + sub;
        nx = l - 2;
        xv = x - ((4 + sub) << (l - 2));
    }
}

// Distance code, extra bits and extra value of a distance 1..32768.
__device__ inline void dist_code(uint32_t d, uint32_t &code, uint32_t &nx, uint32_t &xv)
{
    const uint32_t x = d - 1;
    if (x < 4) {
        code = x;
        nx = 0;
        xv = 0;
    } else {
        const uint32_t l = 31 - __clz(x); // 2..14
        const uint32_t sub = (x >> (l - 1)) & 1;
        code = 2 * l + sub;
        nx = l - 1;
        xv = x - ((2 + sub) << (l - 1));
    }
}

__device__ inline uint32_t reverse_bits(uint32_t v, uint32_t n) { return __brev(v) >> (32 - n); }

// Exclusive prefix over the workgroup; *total: the sum.
__device__ inline uint64_t block_excl(uint64_t v, Lds &s, uint64_t *total)
{
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x / 64;
    uint64_t inc = v;
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t x = __shfl_up(inc, o, 64);
        if ((int)lane >= o) inc += x;
    }
    if (lane == 63) s.scan[w] = inc;
    __syncthreads();
    uint64_t before = 0, all = 0;
    for (uint32_t k = 0; k < NW; ++k) {
        if (k < w) before += s.scan[k];
        all += s.scan[k];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

// Huffman code lengths for f[0..nsym), at most `limit` bits, written to len[].
// Called by every lane of one wave; the serial part runs on lane 0.
__device__ void build_lengths(const uint32_t *f, uint8_t *len, uint32_t nsym, uint32_t limit, HuffWork &w,
                              uint32_t t)
{
    {
        // rank of each used symbol by (frequency, symbol)
        for (uint32_t s = t; s < nsym; s += 64) {
            len[s] = 0;
            if (!f[s]) continue;
            uint32_t r = 0;
            for (uint32_t j = 0; j < nsym; ++j)
                if (f[j] && (f[j] < f[s] || (f[j] == f[s] && j < s))) ++r;
            w.order[r] = (uint16_t)s;
        }
        uint32_t m = 0;
        for (uint32_t s = t; s < nsym; s += 64) m += f[s] ? 1u : 0u;
        for (int o = 32; o > 0; o >>= 1) m += __shfl_xor(m, o, 64);
        if (t == 0) w.m = m;
    }
    WSYNC();
    if (t == 0) {
        const uint32_t m = w.m;
        if (m == 1) {
            len[w.order[0]] = 1;
        } else if (m > 1) {
            // two queues: leaves in ascending frequency, internal nodes as made
            for (uint32_t i = 0; i < m; ++i) w.wf[i] = f[w.order[i]];
            uint32_t il = 0, ii = m, next = m;
            for (uint32_t k = 0; k + 1 < m; ++k) {
                uint32_t pick[2];
                for (int q = 0; q < 2; ++q) {
                    if (il < m && (ii >= next || w.wf[il] <= w.wf[ii])) pick[q] = il++;
                    else pick[q] = ii++;
                }
                w.wf[next] = w.wf[pick[0]] + w.wf[pick[1]];
                w.parent[pick[0]] = (uint16_t)next;
                w.parent[pick[1]] = (uint16_t)next;
                ++next;
            }
            const uint32_t root = 2 * m - 2;
            w.depth[root] = 0;
            for (int32_t i = (int32_t)root - 1; i >= 0; --i) {
                const uint32_t d = w.depth[w.parent[i]] + 1u;
                w.depth[i] = (uint8_t)(d > 255 ? 255 : d);
            }
            // lengths by count, those beyond the limit clamped to it; then,
            // while the Kraft sum exceeds 2^limit, one longest leaf goes and a
            // shorter one splits into two a level down (the sum drops by one
            // unit, the leaf count stays), so the code ends complete; lengths
            // are handed out least frequent first
            for (uint32_t l = 0; l < 16; ++l) w.bl_count[l] = 0;
            for (uint32_t i = 0; i < m; ++i) ++w.bl_count[w.depth[i] > limit ? limit : w.depth[i]];
            uint32_t kraft = 0;
            for (uint32_t l = 1; l <= limit; ++l) kraft += w.bl_count[l] << (limit - l);
            while (kraft > (1u << limit)) {
                uint32_t bits = limit - 1;
                while (bits > 0 && w.bl_count[bits] == 0) --bits;
                if (bits == 0) break; // cannot happen for m <= 2^limit
                --w.bl_count[limit];
                --w.bl_count[bits];
                w.bl_count[bits + 1] += 2;
                --kraft;
            }
            uint32_t i = 0;
            for (uint32_t l = limit; l >= 1; --l)
                for (uint32_t c = w.bl_count[l]; c > 0 && i < m; --c) len[w.order[i++]] = (uint8_t)l;
        }
    }
    WSYNC();
}

// Canonical codes (RFC 1951 3.2.2), stored bit-reversed for an LSB-first
// stream; by lane 0 of the calling wave.
__device__ void assign_codes(const uint8_t *len, uint16_t *code, uint32_t nsym, uint32_t t)
{
    if (t == 0) {
        uint32_t cnt[16] = {}, next[16] = {};
        for (uint32_t s = 0; s < nsym; ++s) ++cnt[len[s]];
        cnt[0] = 0;
        uint32_t c = 0;
        for (uint32_t l = 1; l < 16; ++l) {
            c = (c + cnt[l - 1]) << 1;
            next[l] = c;
        }
        for (uint32_t s = 0; s < nsym; ++s)
            code[s] = len[s] ? (uint16_t)reverse_bits(next[len[s]]++, len[s]) : 0;
    }
    WSYNC();
}

// Bits into the staging area: OR at bit offset pos (atomic when shared).
__device__ inline void put_bits(uint32_t *out, uint64_t pos, uint64_t bits, uint32_t n, bool atomic)
{
    if (n == 0) return;
    const uint32_t w = (uint32_t)(pos >> 5), sh = (uint32_t)(pos & 31);
    const uint64_t lo = bits << sh;          // up to 63 bits used (n <= 48, sh <= 31 -> n + sh <= 79)
    const uint32_t w0 = (uint32_t)lo, w1 = (uint32_t)(lo >> 32);
    const uint32_t w2 = sh ? (uint32_t)(bits >> (64 - sh)) : 0u;
    const uint32_t words = (sh + n + 31) / 32;
    if (atomic) {
        atomicOr(&out[w], w0);
        if (words > 1) atomicOr(&out[w + 1], w1);
        if (words > 2) atomicOr(&out[w + 2], w2);
    } else {
        out[w] |= w0;
        if (words > 1) out[w + 1] |= w1;
        if (words > 2) out[w + 2] |= w2;
    }
}

// The code bits of one symbol: (bits, count).
__device__ inline uint32_t symbol_bits(const Tables &T, const uint8_t *src, const uint16_t *dist, uint32_t sym,
                                       uint64_t *bits)
{
    const uint32_t pos = sym & 0xFFFF;
    if (!(sym & MATCH)) {
        const uint32_t c = src[pos];
        if (bits) *bits = T.c_ll[c];
        return T.l_ll[c];
    }
    const uint32_t L = ((sym >> 16) & 0x7FFF) + 3;
    uint32_t lc, lnx, lxv, dc, dnx, dxv;
    len_code(L, lc, lnx, lxv);
    dist_code(ld_dist(dist, pos), dc, dnx, dxv);
    const uint32_t n0 = T.l_ll[lc], n1 = T.l_d[dc];
    if (bits) {
        uint64_t b = T.c_ll[lc];
        uint32_t at = n0;
        b |= (uint64_t)lxv << at;
        at += lnx;
        b |= (uint64_t)T.c_d[dc] << at;
        at += n1;
        b |= (uint64_t)dxv << at;
        *bits = b;
    }
    return n0 + lnx + n1 + dnx;
}

__device__ inline uint32_t symbol_at(const Lds &s, const uint32_t *syms, uint32_t i)
{
    uint32_t lo = 0, hi = NSEG; // last segment with seg_base <= i
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (s.seg_base[mid] <= i) lo = mid;
        else hi = mid;
    }
    return syms[lo * SEG + (i - s.seg_base[lo])];
}

// One dynamic block's codes, by one wave: the histogram of its symbols
// [b0, b1), length-limited Huffman codes (15 bits; 7 for the code-length
// code), the runs of code lengths its header sends (16/17/18), its header's
// size and the size of its symbols.
__device__ void block_codes(Tables &T, const Lds &s, const uint8_t *src, const uint16_t *dist, const uint32_t *syms,
                            uint32_t b0, uint32_t b1, uint32_t lane)
{
    for (uint32_t i = lane; i < 288; i += 64) T.f_ll[i] = 0;
    for (uint32_t i = lane; i < 32; i += 64) T.f_d[i] = 0;
    for (uint32_t i = lane; i < 19; i += 64) T.f_cl[i] = 0;
    WSYNC();
    for (uint32_t i = b0 + lane; i < b1; i += 64) {
        const uint32_t sym = symbol_at(s, syms, i);
        const uint32_t pos = sym & 0xFFFF;
        if (sym & MATCH) {
            uint32_t lc, nx, xv, dc;
            len_code(((sym >> 16) & 0x7FFF) + 3, lc, nx, xv);
            dist_code(ld_dist(dist, pos), dc, nx, xv);
            atomicAdd(&T.f_ll[lc], 1u);
            atomicAdd(&T.f_d[dc], 1u);
        } else {
            atomicAdd(&T.f_ll[src[pos]], 1u);
        }
    }
    WSYNC();
    if (lane == 0) {
        T.f_ll[256] = 1;
        // two distance codes at least, so the distance code is complete
        uint32_t used = 0;
        for (uint32_t i = 0; i < 30; ++i) used += T.f_d[i] ? 1u : 0u;
        if (used < 2) {
            if (!T.f_d[0]) T.f_d[0] = 1;
            else T.f_d[1] = 1;
        }
    }
    WSYNC();
    build_lengths(T.f_ll, T.l_ll, 286, 15, T.w, lane);
    build_lengths(T.f_d, T.l_d, 30, 15, T.w, lane);
    if (lane == 0) {
        uint32_t hl = 286, hd = 30;
        while (hl > 257 && !T.l_ll[hl - 1]) --hl;
        while (hd > 1 && !T.l_d[hd - 1]) --hd;
        T.hlit = hl;
        T.hdist = hd;
        // runs of code lengths over the HLIT + HDIST lengths
        const uint32_t N = hl + hd;
        auto lenat = [&](uint32_t i) -> uint32_t { return i < hl ? T.l_ll[i] : T.l_d[i - hl]; };
        uint32_t ncl = 0, i = 0;
        while (i < N) {
            const uint32_t v = lenat(i);
            uint32_t rl = 1;
            while (i + rl < N && lenat(i + rl) == v) ++rl;
            if (v == 0) {
                while (rl >= 11) {
                    const uint32_t r = rl < 138 ? rl : 138;
                    T.cl_sym[ncl++] = (uint16_t)(18 | ((r - 11) << 5));
                    rl -= r;
                    i += r;
                }
                if (rl >= 3) {
                    T.cl_sym[ncl++] = (uint16_t)(17 | ((rl - 3) << 5));
                    i += rl;
                    rl = 0;
                }
                for (; rl > 0; --rl, ++i) T.cl_sym[ncl++] = 0;
            } else {
                T.cl_sym[ncl++] = (uint16_t)v;
                ++i;
                --rl;
                while (rl >= 3) {
                    const uint32_t r = rl < 6 ? rl : 6;
                    T.cl_sym[ncl++] = (uint16_t)(16 | ((r - 3) << 5));
                    rl -= r;
                    i += r;
                }
                for (; rl > 0; --rl, ++i) T.cl_sym[ncl++] = (uint16_t)v;
            }
        }
        T.n_cl = ncl;
        for (uint32_t k = 0; k < ncl; ++k) ++T.f_cl[T.cl_sym[k] & 31];
        // decoders take only a complete code-length code: two symbols at least
        uint32_t used = 0;
        for (uint32_t k = 0; k < 19; ++k) used += T.f_cl[k] ? 1u : 0u;
        if (used < 2) {
            if (!T.f_cl[0]) T.f_cl[0] = 1;
            else T.f_cl[1] = 1;
        }
    }
    WSYNC();
    build_lengths(T.f_cl, T.l_cl, 19, 7, T.w, lane);
    assign_codes(T.l_ll, T.c_ll, 286, lane);
    assign_codes(T.l_d, T.c_d, 30, lane);
    assign_codes(T.l_cl, T.c_cl, 19, lane);
    uint32_t part = 0;
    for (uint32_t i = b0 + lane; i < b1; i += 64) part += symbol_bits(T, src, dist, symbol_at(s, syms, i), nullptr);
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o, 64);
    if (lane == 0) {
        uint32_t hc = 19;
        while (hc > 4 && !T.l_cl[CL_ORDER[hc - 1]]) --hc;
        T.hclen = hc;
        uint32_t bits = 3 + 5 + 5 + 4 + 3 * hc;
        for (uint32_t k = 0; k < T.n_cl; ++k) {
            const uint32_t c = T.cl_sym[k] & 31;
            bits += T.l_cl[c] + (c == 16 ? 2u : c == 17 ? 3u : c == 18 ? 7u : 0u);
        }
        T.header_bits = bits;
        T.sym_bits = part;
    }
    WSYNC();
}

__device__ inline uint8_t framed_byte(const Lds &s, const uint8_t *src, uint32_t j, uint32_t body_len,
                                      uint32_t total, uint32_t crc)
{
    const uint32_t n = s.n;
    if (j < 18) {
        // ID1 ID2 CM=8 FLG=FEXTRA, MTIME 0, XFL 0, OS 255, XLEN 6, 'B' 'C' SLEN 2, BC
        const uint32_t bc = total - 1;
        switch (j) {
        case 0: return 0x1f;
        case 1: return 0x8b;
        case 2: return 8;
        case 3: return 4;
        case 9: return 0xff;
        case 10: return 6;
        case 12: return 'B';
        case 13: return 'C';
        case 14: return 2;
        case 16: return (uint8_t)(bc & 0xff);
        case 17: return (uint8_t)(bc >> 8);
        default: return 0;
        }
    }
    const uint32_t k = j - 18;
    if (k < body_len) {
        if (s.framing == FRAMING_EMPTY) return k == 0 ? 0x03 : 0x00;
        if (s.framing == FRAMING_STORED) {
            if (k == 0) return 0x01;
            if (k == 1) return (uint8_t)(n & 0xff);
            if (k == 2) return (uint8_t)(n >> 8);
            if (k == 3) return (uint8_t)(~n & 0xff);
            if (k == 4) return (uint8_t)((~n >> 8) & 0xff);
            return src[k - 5];
        }
        return (uint8_t)(s.b.e.out[k >> 2] >> (8 * (k & 3)));
    }
    const uint32_t q = k - body_len;
    const uint32_t v = q < 4 ? crc : n;
    return (uint8_t)(v >> (8 * (q & 3)));
}

__global__ __launch_bounds__(DT) void deflate_kernel(DeflateArgs a)
{
    __shared__ Lds s;
    const uint32_t t = threadIdx.x;
    uint16_t *dist = a.dist + (uint64_t)blockIdx.x * 65536;
    uint32_t *syms = a.syms + (uint64_t)blockIdx.x * 65536;
    for (;;) {
        if (t == 0) s.chunk = atomicAdd(a.queue, 1u);
        __syncthreads();
        const uint32_t chunk = s.chunk;
        if (chunk >= a.n) break; // every wave of every workgroup leaves here
        const uint64_t b0 = a.bounds[chunk];
        const uint32_t n = (uint32_t)(a.bounds[chunk + 1] - b0);
        if (t == 0) s.n = n;
        const uint8_t *src = a.in + b0;
        uint8_t *lens_g = (uint8_t *)syms; // match lengths pass through the symbol scratch
        uint8_t *D = s.b.m.data;
        {
            // the chunk into LDS, dword by dword from the aligned global
            // dwords that hold it, zero past its end
            uint32_t *D32 = (uint32_t *)D;
            for (uint32_t k = t; k < (65536 + 64) / 4; k += DT) {
                uint32_t v = 0;
                const uint32_t off = 4 * k;
                if (off < n) {
                    const uint8_t *q = src + off;
                    const uint32_t al = (uint32_t)((uintptr_t)q & 3);
                    const uint32_t *w = (const uint32_t *)(q - al);
                    const uint32_t lo = w[0];
                    const uint32_t hi = al && off + 4 - al < n ? w[1] : 0u;
                    v = __builtin_amdgcn_alignbyte(hi, lo, al);
                    const uint32_t keep = n - off;
                    if (keep < 4) v &= (1u << (8 * keep)) - 1;
                }
                D32[k] = v;
            }
        }
        for (uint32_t i = t; i < NBUCKET; i += DT) {
            s.c.ch.head[i] = 0;
            s.c.ch.stamp[i] = 0;
        }
        __syncthreads();
        auto ring_at = [](uint32_t q) -> uint32_t { return q >= RING ? q - RING : q; };

        // 1. Rounds of ROUND positions.
        // Hash chains, in sub-rounds of 512 positions, one per thread: each
        // position first links to the last position of an earlier sub-round
        // with its hash; then the sub-round's positions of one hash are
        // ranked from the highest down by atomicMax over stamps that grow
        // with sub-round and iteration, and each winner links the previous
        // one to itself, up to KLINK per hash (positions beyond stay out of
        // the chain). No outcome depends on thread order. A position's link
        // lives at p % RING: the ring holds the window behind the round and
        // the round, and no walk reads a link farther back than the window.
        // Matches: then each thread takes 4 consecutive positions of the
        // round, all bytes and candidates read from the chunk in LDS. A
        // position starts from the previous one's match, one byte on (and
        // searches no further when that is already LAZY_MAX long: the lazy
        // parse then looks no further either); then distances 1 to 4, then
        // the chain, KB links at a time. A link is taken only when longer
        // than the best so far (as in zlib's longest_match), so it first
        // shows the 4 bytes ending at that length, and only links whose
        // bytes there equal the position's own are compared in full, 16
        // bytes at a time.
        for (uint32_t r0 = 0; r0 < n; r0 += ROUND) {
            for (uint32_t sr = r0; sr < r0 + ROUND && sr < n; sr += DT) {
                const uint32_t p = sr + t;
                const uint32_t ridx = sr / DT;
                bool active = p + 2 < n;
                uint32_t h = 0;
                if (active) {
                    h = hash3(D + p);
                    const uint32_t ho = s.c.ch.head[h];
                    s.b.m.prev[ring_at(p)] = ho ? (uint16_t)(ho - 1) : NO_POS;
                }
                for (uint32_t k = 0; k < KLINK; ++k) {
                    const uint32_t before = active ? s.c.ch.stamp[h] : 0u;
                    __syncthreads();
                    const uint32_t stamp = ((ridx * KLINK + k + 1) << 16) | (p - sr);
                    if (active) atomicMax(&s.c.ch.stamp[h], stamp);
                    __syncthreads();
                    if (active && s.c.ch.stamp[h] == stamp) {
                        if (k == 0) s.c.ch.head[h] = (uint16_t)(p + 1); // the sub-round's last becomes the head
                        else if ((before >> 16) == ridx * KLINK + k) s.b.m.prev[ring_at(sr + (before & 0xFFFF))] = (uint16_t)p;
                        active = false;
                    }
                    if (!__syncthreads_or(active)) break;
                }
            }
            // Pre-pass: each thread's 4 consecutive positions, from the
            // previous position's match one byte on, then distances 1 to 4.
            // A length of LAZY_MAX or more is final (the lazy parse looks no
            // further); other positions go on to the chain pass.
            uint32_t cl = 0, cd = 0; // the previous position's match
            for (uint32_t j = 0; j < 4; ++j) {
                const uint32_t p = r0 + 4 * t + j;
                if (p >= n) break;
                const uint32_t maxlen = n - p < MAX_LEN ? n - p : MAX_LEN;
                uint32_t best = 0, bd = 0;
                if (cl > 3) {
                    best = cl - 1 < maxlen ? cl - 1 : maxlen;
                    bd = cd;
                }
                const bool chain = p + 2 < n && best < LAZY_MAX;
                if (chain) {
                    const uint4 P = lds16(D, p);
#pragma unroll
                    for (uint32_t d = 1; d <= 4; ++d) {
                        if (d <= p) {
                            uint32_t L = prefix16(P, lds16(D, p - d));
                            for (uint32_t e = 16; L == e && e < maxlen; e += 16)
                                L = e + prefix16(lds16(D, p + e), lds16(D, p - d + e));
                            L = L < maxlen ? L : maxlen;
                            if (L > best || (L == best && d < bd)) {
                                best = L;
                                bd = d;
                            }
                        }
                    }
                }
                if (best < 3 || (!chain && best == 3 && bd > TOO_FAR)) best = 0;
                lens_g[p] = (uint8_t)(best ? best - 2 : 0);
                dist[p] = (uint16_t)bd;
                cl = best;
                cd = bd;
            }
            if (t == 0) s.qnext = 0;
            __syncthreads();
            // Chain pass: each lane takes the round's positions one at a time
            // from a queue, so lanes whose chains end early take more. A
            // position starts from its pre-pass match (read back past the L1:
            // another wave wrote it) and walks the chain KB links at a time. A
            // link is taken only when longer than the best so far (as in
            // zlib's longest_match), so it first shows the 4 bytes ending at
            // that length; only links whose bytes there equal the position's
            // own are compared in full, 16 bytes at a time. The walk issues
            // each link's loads without a branch (a link past the chain's end
            // reads in-bounds slots of the position itself) and waits only for
            // the next link; the compares come after it.
            {
                constexpr uint32_t KB = 8;
                const uint32_t cnt = n - r0 < ROUND ? n - r0 : ROUND;
                uint32_t p = 0, q = NO_POS, k = 0, best = 0, bd = 0, maxlen = 0;
                uint4 P = make_uint4(0, 0, 0, 0);
                bool have = false;
                for (;;) {
                    if (!have) {
                        const uint32_t idx = atomicAdd(&s.qnext, 1u);
                        if (idx >= cnt) break;
                        p = r0 + idx;
                        if (p + 2 >= n) continue;
                        const uint32_t lw = __hip_atomic_load((const uint32_t *)(lens_g + (p & ~3u)), __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT);
                        const uint32_t l8 = (lw >> (8 * (p & 3))) & 0xFFu;
                        best = l8 ? l8 + 2 : 0;
                        if (best >= LAZY_MAX) continue;
                        bd = ld_dist(dist, p);
                        maxlen = n - p < MAX_LEN ? n - p : MAX_LEN;
                        P = lds16(D, p);
                        q = s.b.m.prev[ring_at(p)];
                        k = 0;
                        have = true;
                    }
                    const bool walk = q != NO_POS && p - q <= WINDOW && k < CHAIN && best < NICE && best < maxlen;
                    if (walk) {
                        // the bytes a link must share to be longer than b0:
                        // [b0 - 3, b0] from b0 = 3 on, the first 3 below it
                        const uint32_t b0 = best;
                        const uint32_t at = b0 >= 3 ? b0 - 3 : 0;
                        const uint32_t mask = b0 >= 3 ? ~0u : 0xFFFFFFu;
                        const uint32_t pw = lds4(D, p + at) & mask;
                        uint32_t c[KB], w0[KB], w1[KB];
#pragma unroll
                        for (uint32_t i = 0; i < KB; ++i) {
                            const bool ok = q != NO_POS && p - q <= WINDOW && k < CHAIN;
                            const uint32_t qq = ok ? q : p;
                            c[i] = ok ? q : NO_POS;
                            const uint32_t *wp = (const uint32_t *)(D + ((qq + at) & ~3u));
                            w0[i] = wp[0];
                            w1[i] = wp[1];
                            const uint32_t nq = s.b.m.prev[ring_at(qq)];
                            q = ok ? nq : q;
                            k += ok ? 1u : 0u;
                        }
#pragma unroll
                        for (uint32_t i = 0; i < KB; ++i) {
                            const uint32_t cw = __builtin_amdgcn_alignbyte(w1[i], w0[i], (c[i] + at) & 3) & mask;
                            if (c[i] != NO_POS && cw == pw) {
                                const uint32_t cq = c[i];
                                uint32_t L = prefix16(P, lds16(D, cq));
                                // a 16-byte prefix that matches in full goes on 16 bytes at a time
                                for (uint32_t e = 16; L == e && e < maxlen; e += 16)
                                    L = e + prefix16(lds16(D, p + e), lds16(D, cq + e));
                                L = L < maxlen ? L : maxlen;
                                const uint32_t d = p - cq;
                                if (L > best || (L == best && d < bd)) {
                                    best = L;
                                    bd = d;
                                }
                            }
                        }
                    }
                    if (!(q != NO_POS && p - q <= WINDOW && k < CHAIN && best < NICE && best < maxlen)) {
                        if (best < 3 || (best == 3 && bd > TOO_FAR)) best = 0;
                        lens_g[p] = (uint8_t)(best ? best - 2 : 0);
                        dist[p] = (uint16_t)bd;
                        have = false;
                    }
                }
            }
            __syncthreads(); // before the next round rewrites the hash heads and ring slots
        }
        // the lengths into LDS, over the chains (coherent loads: written by
        // other threads of this workgroup in this launch)
        for (uint32_t i = t; i < (n + 3) / 4; i += DT)
            ((uint32_t *)s.b.e.lens)[i] = __hip_atomic_load((const uint32_t *)lens_g + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();

        // 2. lazy parse, one thread per 2 KiB segment
        if (t < NSEG) {
            uint32_t p = t * SEG, k = 0;
            const uint32_t end = p + SEG < n ? p + SEG : n;
            while (p < end) {
                uint32_t L = s.b.e.lens[p] ? s.b.e.lens[p] + 2u : 0u;
                if (L > end - p) L = end - p;
                if (L < 3) L = 0;
                if (L && L < LAZY_MAX && p + 1 < end) {
                    uint32_t L1 = s.b.e.lens[p + 1] ? s.b.e.lens[p + 1] + 2u : 0u;
                    if (L1 > end - p - 1) L1 = end - p - 1;
                    if (L1 > L) L = 0;
                }
                if (L) {
                    syms[t * SEG + k++] = MATCH | ((L - 3) << 16) | p;
                    p += L;
                } else {
                    syms[t * SEG + k++] = p;
                    p += 1;
                }
            }
            s.seg_count[t] = t * SEG < n ? k : 0;
        }
        __syncthreads();
        if (t == 0) {
            uint32_t run = 0;
            for (uint32_t g = 0; g < NSEG; ++g) {
                s.seg_base[g] = run;
                run += s.seg_count[g];
            }
            s.seg_base[NSEG] = run;
        }
        __syncthreads();
        const uint32_t nsyms = s.seg_base[NSEG];

        // 3. codes: each dynamic block of BSYM symbols gets its own codes,
        // built by one wave per block, NW blocks at a time, in the lengths'
        // space (free after the parse); 4. emit: the blocks back to back
        // into the staging area, each block's symbols placed by a workgroup
        // prefix sum of their bit counts. A chunk whose blocks would not beat
        // storing it (or would not fit a block) is stored.
        for (uint32_t i = t; i < 16384; i += DT) s.b.e.out[i] = 0;
        if (t == 0) s.total_bits = 0;
        const uint64_t cap_bits = n ? 8ull * (n + 5 < DEFLATE_MAX_STREAM ? n + 5 : DEFLATE_MAX_STREAM) : 0;
        bool overflow = false;
        uint64_t run = 0;
        Tables *tb = reinterpret_cast<Tables *>(s.b.e.lens);
        const uint32_t nblk = (nsyms + BSYM - 1) / BSYM;
        for (uint32_t g = 0; g < nblk && !overflow; g += NW) {
            {
                const uint32_t w = t / 64;
                if (g + w < nblk) {
                    const uint32_t b0 = (g + w) * BSYM, b1 = b0 + BSYM < nsyms ? b0 + BSYM : nsyms;
                    block_codes(tb[w], s, src, dist, syms, b0, b1, t & 63);
                }
            }
            __syncthreads();
            for (uint32_t j = 0; j < NW && g + j < nblk; ++j) {
                const Tables &T = tb[j];
                const uint32_t b0 = (g + j) * BSYM, b1 = b0 + BSYM < nsyms ? b0 + BSYM : nsyms;
                const uint64_t block_bits = (uint64_t)T.header_bits + T.sym_bits + T.l_ll[256];
                if (run + block_bits > cap_bits) {
                    overflow = true; // the same on every thread
                    break;
                }
                if (t == 0) {
                    uint64_t at = run;
                    auto put = [&](uint64_t v, uint32_t nb) {
                        put_bits(s.b.e.out, at, v, nb, false);
                        at += nb;
                    };
                    put(b1 == nsyms ? 1u : 0u, 1); // BFINAL on the last block
                    put(2, 2);                     // BTYPE 10: dynamic codes
                    put(T.hlit - 257, 5);
                    put(T.hdist - 1, 5);
                    put(T.hclen - 4, 4);
                    for (uint32_t k = 0; k < T.hclen; ++k) put(T.l_cl[CL_ORDER[k]], 3);
                    for (uint32_t k = 0; k < T.n_cl; ++k) {
                        const uint32_t c = T.cl_sym[k] & 31, x = T.cl_sym[k] >> 5;
                        put(T.c_cl[c], T.l_cl[c]);
                        if (c == 16) put(x, 2);
                        else if (c == 17) put(x, 3);
                        else if (c == 18) put(x, 7);
                    }
                }
                __syncthreads();
                uint64_t at0 = run + T.header_bits;
                for (uint32_t i0 = b0; i0 < b1; i0 += DT) {
                    const uint32_t i = i0 + t;
                    uint64_t bits = 0;
                    uint32_t nb = 0;
                    if (i < b1) nb = symbol_bits(T, src, dist, symbol_at(s, syms, i), &bits);
                    uint64_t all;
                    const uint64_t at = at0 + block_excl(nb, s, &all);
                    if (nb) put_bits(s.b.e.out, at, bits, nb, true);
                    at0 += all;
                }
                if (t == 0) put_bits(s.b.e.out, at0, T.c_ll[256], T.l_ll[256], true);
                run += block_bits;
                __syncthreads();
            }
            __syncthreads(); // before the next group's codes replace these tables
        }
        if (t == 0) {
            s.total_bits = run;
            s.framing = n == 0 ? FRAMING_EMPTY : overflow ? FRAMING_STORED : FRAMING_DEFLATED;
        }
        __syncthreads();

        // 5. the framed block into the chunk's slot
        const uint32_t body = s.framing == FRAMING_EMPTY    ? 2u
                              : s.framing == FRAMING_STORED ? n + 5
                                                            : (uint32_t)((s.total_bits + 7) / 8);
        const uint32_t total = 18 + body + 8;
        const uint32_t crc = a.crc[chunk];
        uint8_t *slot = a.slots + (uint64_t)chunk * DEFLATE_SLOT;
        for (uint32_t j = t; j < total; j += DT) slot[j] = framed_byte(s, src, j, body, total, crc);
        if (t == 0) {
            a.block_size[chunk] = total;
            a.framing[chunk] = (uint8_t)s.framing;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void gather_blocks_kernel(const uint8_t *slots, const uint32_t *block_size,
                                                            const uint64_t *offsets, uint32_t n, uint8_t *out)
{
    for (uint32_t b = blockIdx.x; b < n; b += gridDim.x) {
        const uint8_t *src = slots + (uint64_t)b * DEFLATE_SLOT;
        uint8_t *dst = out + offsets[b];
        const uint32_t len = block_size[b];
        for (uint32_t j = threadIdx.x; j < len; j += 256) dst[j] = src[j];
    }
}

} // namespace

hipError_t launch_deflate(const DeflateArgs &a, uint32_t workgroups, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(deflate_kernel, dim3(workgroups), dim3(DT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_gather_blocks(const uint8_t *slots, const uint32_t *block_size, const uint64_t *offsets, uint32_t n,
                                uint8_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    const uint32_t grid = n < 4096 ? n : 4096;
    hipLaunchKernelGGL(gather_blocks_kernel, dim3(grid), dim3(256), 0, s, slots, block_size, offsets, n, out);
    return hipGetLastError();
}

} // namespace frhip
