// bam.hip -- BAM record boundaries and SoA column decode over an inflated batch.
//
// Result semantics follow the reference's device scan
// (crates/fritillaria-bam/src/columnar/blocked.rs:106-282 and
//  crates/fritillaria-cuda/kernels/bam_decode.cu:63-291): the true record chain
// from `start`, `block_size >= 32` and `fields_fit` on every record reached from
// a proven boundary, a tail at the first record the buffer cuts short, and a
// malformed record reported at the chain's entry into the member holding it.
//
// The speculation is ours. The reference guesses that each member starts on a
// record, which never holds for noodles-written files (it cuts members every
// 65,495 bytes whatever the records). Here every byte offset in the first
// `window` bytes of a member is a candidate: a wave tests 64 at a time with a
// short walk to the window's end, the survivors must agree on where they land,
// and one exact walk from that landing to the member end records the offsets.
// A later reconcile adopts a member's walk only when the true chain enters it at
// one of its survivors, so a wrong guess costs a fallback walk, never a wrong
// answer.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frhip_internal.h"
#include "pipeline_internal.h"

namespace frhip {
namespace {

enum : uint64_t { BAM_OK = 0, BAM_MALFORMED = 1 };
// Members the parallel reconcile looks back over for the closest one holding a
// record start (runs of members inside one long record); longer runs take the
// serial reconcile.
constexpr uint32_t FAST_MAX_HOPS = 4096;

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v)
{
    // the builtin returns int: each half is narrowed to uint32_t before it is
    // widened, or the low half would be sign-extended over the high one
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint32_t ld32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}
__device__ __forceinline__ uint32_t ld16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }

// Four little-endian words at a 4-byte aligned address: one 16-byte load.
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4))) __attribute__((aligned(4)));

// A record's 36 fixed bytes at cur (cur + 36 <= len) as nine little-endian
// words w[k] = bytes 4k..4k+3: records lie at any byte offset, so three
// 16-byte loads from the aligned word at or below cur, funnel-shifted (one
// load per byte would touch a cache line per byte where records are far
// apart, and costs ~36 instructions anyway); within 48 bytes of len, bytes.
__device__ __forceinline__ void fixed_words(const uint8_t *buf, uint64_t len, uint64_t cur, uint32_t *w)
{
    const uint64_t a4 = cur & ~3ull;
    if (a4 + 48 <= len) {
        const u32x4_a4 *q = reinterpret_cast<const u32x4_a4 *>(buf + a4);
        const u32x4_a4 v0 = q[0], v1 = q[1], v2 = q[2];
        const uint32_t d[12] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
        const uint32_t sh = (uint32_t)(cur & 3) * 8;
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = __builtin_amdgcn_alignbit(d[k + 1], d[k], sh);
    } else {
        const uint8_t *p = buf + cur;
        uint32_t h[36];
#pragma unroll
        for (int i = 0; i < 36; ++i) h[i] = p[i];
#pragma unroll
        for (int k = 0; k < 9; ++k) w[k] = h[4 * k] | h[4 * k + 1] << 8 | h[4 * k + 2] << 16 | h[4 * k + 3] << 24;
    }
}

// One step of the chain from `cur` (a record start candidate); the walk stops
// at `until`. Returns 0 advanced, 1 reached until, 2 ran out of buffer
// (partial record: the tail), 3 malformed.
struct Step {
    uint64_t next;
    int kind;
};

__device__ __forceinline__ Step step(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until)
{
    if (cur >= until) return {cur, 1};
    const uint64_t a4 = cur & ~3ull;
    if (a4 + 32 <= len) {
        // the fields the step reads: two 16-byte loads from the aligned word
        // at or below the record, funnel-shifted (one wait per record, and a
        // walk over members far apart touches two lines, not one per byte)
        const u32x4_a4 *q = reinterpret_cast<const u32x4_a4 *>(buf + a4);
        const u32x4_a4 v0 = q[0], v1 = q[1];
        const uint32_t sh = (uint32_t)(cur & 3) * 8;
        const uint32_t bs = __builtin_amdgcn_alignbit(v0.y, v0.x, sh);
        if (bs < 32) return {cur, 3};
        const uint64_t end = cur + 4 + (uint64_t)bs;
        if (end > len) return {cur, 2};
        const uint32_t name_len = __builtin_amdgcn_alignbit(v1.x, v0.w, sh) & 0xFF;
        const uint64_t n_cigar = __builtin_amdgcn_alignbit(v1.y, v1.x, sh) & 0xFFFF;
        const uint64_t l_seq = __builtin_amdgcn_alignbit(v1.z, v1.y, sh);
        const uint64_t need = 36ull + name_len + 4ull * n_cigar + (l_seq + 1) / 2 + l_seq;
        if (name_len == 0 || cur + need > end) return {cur, 3};
        return {end, 0};
    }
    if (cur + 24 <= len) {
        // the fields the step reads, loaded together (one wait per record)
        const uint8_t *r = buf + cur;
        uint32_t h[24];
#pragma unroll
        for (int i = 0; i < 24; ++i) h[i] = r[i];
        asm volatile("" ::"v"(h[0]), "v"(h[1]), "v"(h[2]), "v"(h[3]), "v"(h[12]), "v"(h[16]), "v"(h[17]),
                     "v"(h[20]), "v"(h[21]), "v"(h[22]), "v"(h[23]));
        const uint32_t bs = h[0] | h[1] << 8 | h[2] << 16 | h[3] << 24;
        if (bs < 32) return {cur, 3};
        const uint64_t end = cur + 4 + (uint64_t)bs;
        if (end > len) return {cur, 2};
        const uint32_t name_len = h[12];
        const uint64_t n_cigar = h[16] | h[17] << 8;
        const uint64_t l_seq = h[20] | h[21] << 8 | h[22] << 16 | h[23] << 24;
        const uint64_t need = 36ull + name_len + 4ull * n_cigar + (l_seq + 1) / 2 + l_seq;
        if (name_len == 0 || cur + need > end) return {cur, 3};
        return {end, 0};
    }
    if (cur + 4 > len) return {cur, 2};
    uint32_t bs = ld32(buf + cur);
    if (bs < 32) return {cur, 3};
    uint64_t end = cur + 4 + (uint64_t)bs;
    if (end > len) return {cur, 2};
    uint32_t name_len = buf[cur + 12];
    uint64_t n_cigar = ld16(buf + cur + 16);
    uint64_t l_seq = ld32(buf + cur + 20);
    uint64_t need = 36ull + name_len + 4ull * n_cigar + (l_seq + 1) / 2 + l_seq;
    if (name_len == 0 || cur + need > end) return {cur, 3};
    return {end, 0};
}

// Candidate filter for the speculation only (never for the true chain): a
// record that cannot be whole inside the buffer is dropped rather than treated
// as a tail, and the fixed fields must look like a record -- IDs and positions
// at least -1, a name of at least one printable SAM QNAME character followed by
// its NUL. A true record that fails this only costs a fallback walk.
__device__ __forceinline__ Step cand_step(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until)
{
    if (cur >= until) return {cur, 1};
    if (cur + 36 > len) return {cur, 3};
    // the 36 fixed bytes, loaded together (every check fails alike, so none
    // needs the others' loads to wait)
    const uint8_t *r = buf + cur;
    uint32_t w[9];
    fixed_words(buf, len, cur, w);
    const uint32_t bs = w[0];
    if (bs < 32) return {cur, 3};
    const uint64_t end = cur + 4 + (uint64_t)bs;
    if (end > len) return {cur, 3};
    const uint32_t name_len = w[3] & 0xFF;
    const uint64_t n_cigar = w[4] & 0xFFFF;
    const uint64_t l_seq = w[5];
    const uint64_t need = 36ull + name_len + 4ull * n_cigar + (l_seq + 1) / 2 + l_seq;
    if (name_len == 0 || cur + need > end) return {cur, 3};
    if ((int32_t)w[1] < -1 || (int32_t)w[2] < -1 || (int32_t)w[6] < -1 || (int32_t)w[7] < -1)
        return {cur, 3};
    // a QNAME is at least one printable character, then the NUL (all inside
    // the record, so inside the buffer)
    if (name_len < 2 || r[35 + name_len] != 0) return {cur, 3};
    const uint32_t n = name_len - 1;
    for (uint32_t k = 0; k < n; k += 16) {
        bool bad = false;
        if (cur + 36 + k + 16 <= len) {
            uint32_t c[16];
#pragma unroll
            for (int i = 0; i < 16; ++i) c[i] = r[36 + k + i];
            asm volatile("" ::"v"(c[0]), "v"(c[1]), "v"(c[2]), "v"(c[3]), "v"(c[4]), "v"(c[5]), "v"(c[6]), "v"(c[7]),
                         "v"(c[8]), "v"(c[9]), "v"(c[10]), "v"(c[11]), "v"(c[12]), "v"(c[13]), "v"(c[14]), "v"(c[15]));
#pragma unroll
            for (uint32_t i = 0; i < 16; ++i) bad |= k + i < n && (c[i] < 33 || c[i] > 126);
        } else {
            for (uint32_t i = k; i < n && i < k + 16; ++i) {
                const uint32_t ch = r[36 + i];
                bad |= ch < 33 || ch > 126;
            }
        }
        if (bad) return {cur, 3};
    }
    return {end, 0};
}

// Whether a candidate's landing past the window can start a record: the
// fixed fields of a BAM record there (block_size at least 32, IDs and
// positions at least -1, a name of at least one character and its NUL). A
// landing within 36 bytes of the buffer end, or whose name runs past it,
// passes (the tail is the reconcile's business). Checked only where a walk
// leaves its window: a true record's successor always passes, while a chance
// match in random bytes (quality strings, kinetics arrays) almost never does,
// so members a long record covers keep no false survivors.
__device__ __forceinline__ bool landing_plausible(const uint8_t *buf, uint64_t len, uint64_t at)
{
    if (at + 36 > len) return true;
    uint32_t w[9];
    fixed_words(buf, len, at, w);
    const uint32_t name_len = w[3] & 0xFF;
    if (w[0] < 32 || name_len < 2) return false;
    if ((int32_t)w[1] < -1 || (int32_t)w[2] < -1 || (int32_t)w[6] < -1 || (int32_t)w[7] < -1) return false;
    return at + 36 + name_len > len || buf[at + 35 + name_len] == 0;
}

// The first position in [from, to) whose fixed fields could begin a record
// (block_size at least 32 and inside the buffer, IDs and positions at least -1,
// a name of at least two bytes ending in its NUL, the fields fitting the
// block): every record
// that passes cand_step passes this, so no candidate before it can survive.
// A wave tests 1,024 positions per step, each lane 16 consecutive ones from
// three aligned 16-byte loads; within 48 bytes of the buffer end every
// position counts (the candidate walk decides there). Returns `to` when none.
__device__ uint64_t skim_bam(const uint8_t *buf, uint64_t len, uint64_t from, uint64_t to, uint32_t lane)
{
    for (uint64_t c = from - (((uintptr_t)(buf + from)) & 15); c < to; c += 1024) {
        const uint64_t p0 = c + 16ull * lane;
        uint32_t mask = 0;
        if (p0 + 48 <= len) {
            const uint4 *q = reinterpret_cast<const uint4 *>(buf + p0);
            const uint4 v0 = q[0], v1 = q[1], v2 = q[2];
            const uint32_t d[12] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
#pragma unroll
            for (uint32_t o = 0; o < 16; ++o) {
                uint32_t w[8];
#pragma unroll
                for (uint32_t k = 0; k < 8; ++k)
                    w[k] = (o & 3) ? __builtin_amdgcn_alignbit(d[(o >> 2) + k + 1], d[(o >> 2) + k], (o & 3) * 8)
                                   : d[(o >> 2) + k];
                const uint64_t p = p0 + o;
                const uint32_t name_len = w[3] & 0xFF;
                const uint64_t need = 36ull + name_len + 4ull * (w[4] & 0xFFFF) + ((uint64_t)w[5] + 1) / 2 + w[5];
                const bool ok = p >= from && p < to && w[0] >= 32 && p + 4 + (uint64_t)w[0] <= len &&
                                (int32_t)w[1] >= -1 && (int32_t)w[2] >= -1 && (int32_t)w[6] >= -1 &&
                                (int32_t)w[7] >= -1 && name_len >= 2 && need <= 4ull + w[0];
                mask |= (uint32_t)ok << o;
            }
            // and the name's NUL (one byte load for the ~2% of random
            // positions that get this far)
            for (uint32_t m = mask; m; m &= m - 1) {
                const uint32_t o = (uint32_t)__builtin_ctz(m);
                const uint64_t p = p0 + o;
                const uint64_t at = p + 35 + buf[p + 12];
                if (at < len && buf[at] != 0) mask &= ~(1u << o);
            }
        } else {
            for (uint32_t o = 0; o < 16; ++o) mask |= (uint32_t)(p0 + o >= from && p0 + o < to) << o;
        }
        const uint64_t bal = __ballot(mask != 0);
        if (bal) {
            const uint32_t l = (uint32_t)__ffsll((unsigned long long)bal) - 1;
            const uint32_t mk = (uint32_t)__shfl((int)mask, (int)l, 64);
            return c + 16ull * l + (uint32_t)__builtin_ctz(mk);
        }
    }
    return to;
}

// BCF: one step of scan_records (crates/fritillaria-bcf/src/columnar/record.rs:60-95):
// fewer than 8 bytes or a record past the buffer is the tail, l_shared < 24 is
// malformed.
__device__ __forceinline__ Step bcf_step(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until)
{
    if (cur >= until) return {cur, 1};
    if (cur + 8 > len) return {cur, 2};
    const uint32_t ls = ld32(buf + cur), li = ld32(buf + cur + 4);
    if (ls < 24) return {cur, 3};
    const uint64_t end = cur + 8 + (uint64_t)ls + li;
    if (end > len) return {cur, 2};
    return {end, 0};
}

// BCF candidate filter: looks_like_a_record (record.rs:96-160) with the
// header's sample and contig counts; records not whole in the buffer drop.
__device__ __forceinline__ Step bcf_cand_step(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until,
                                              uint32_t samples, uint32_t contigs)
{
    if (cur >= until) return {cur, 1};
    if (cur + 32 > len) return {cur, 3};
    const uint32_t ls = ld32(buf + cur), li = ld32(buf + cur + 4);
    if (ls < 24) return {cur, 3};
    const uint64_t end = cur + 8 + (uint64_t)ls + li;
    if (end > len) return {cur, 3};
    const int32_t chrom = (int32_t)ld32(buf + cur + 8), pos = (int32_t)ld32(buf + cur + 12);
    const int32_t rlen = (int32_t)ld32(buf + cur + 16);
    if (chrom < 0 || (uint32_t)chrom >= contigs || pos < -1 || rlen < 0) return {cur, 3};
    const uint32_t n_allele = ld16(buf + cur + 26);
    const uint32_t n_sample = ld16(buf + cur + 28) | ((uint32_t)buf[cur + 30] << 16);
    const uint32_t n_fmt = buf[cur + 31];
    if (n_sample != samples || n_allele == 0) return {cur, 3};
    if (n_fmt == 0 ? li != 0 : (uint64_t)li < 2ull * n_fmt) return {cur, 3};
    return {end, 0};
}

// The chain's exact step and candidate filter, and the record length the emit
// walk adds, per format.
template <uint32_t F>
__device__ __forceinline__ Step chain_step(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until)
{
    return F == CHAIN_BCF ? bcf_step(buf, len, cur, until) : step(buf, len, cur, until);
}
template <uint32_t F>
__device__ __forceinline__ Step chain_cand(const uint8_t *buf, uint64_t len, uint64_t cur, uint64_t until,
                                           uint32_t samples, uint32_t contigs)
{
    return F == CHAIN_BCF ? bcf_cand_step(buf, len, cur, until, samples, contigs) : cand_step(buf, len, cur, until);
}
template <uint32_t F>
__device__ __forceinline__ uint64_t chain_next(const uint8_t *buf, uint64_t cur)
{
    return F == CHAIN_BCF ? cur + 8 + (uint64_t)ld32(buf + cur) + ld32(buf + cur + 4)
                          : cur + 4 + (uint64_t)ld32(buf + cur);
}

struct Walk {
    uint64_t land;
    uint32_t count;
    int complete;
    int valid;
};

// walk_segment (blocked.rs:141-187), uniform.
template <uint32_t F>
__device__ Walk walk_uniform(const uint8_t *buf, uint64_t len, uint64_t from, uint64_t until)
{
    Walk w{from, 0, 1, 1};
    uint64_t cur = from;
    for (;;) {
        Step s = chain_step<F>(buf, len, cur, until);
        if (s.kind == 1) { w.land = cur; return w; }
        if (s.kind == 2) { w.land = cur; w.complete = 0; return w; }
        if (s.kind == 3) { w.land = cur; w.valid = 0; return w; }
        w.count++;
        cur = s.next;
    }
}

template <uint32_t F>
__global__ __launch_bounds__(64) void bam_spec_kernel(BamSpecArgs a)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint64_t sb = a.block_starts[b];
    const uint64_t sb1 = a.mem_until ? a.mem_until[b] : a.block_starts[b + 1];
    const uint64_t len = a.mem_len ? a.mem_len[b] : a.len;
    if (sb >= sb1) {
        if (lane == 0) {
            a.spec_ok[b] = 0;
            a.n_surv[b] = 0;
        }
        return;
    }
    // candidate walks to the window end; a window without survivors grows by
    // another window, up to the member end, so that a member left with none
    // holds no record start at all (the true start passes the filter): the
    // chain then passes it by, which the parallel reconcile relies on
    uint64_t T = (sb1 - sb) > a.window ? sb + a.window : sb1;
    uint32_t n_surv = 0;
    int agree = 1;
    uint64_t L = 0;
    int Lcomplete = 1;
    const uint64_t lt = (1ull << lane) - 1;
    for (uint64_t base = sb; base < T; base += 64) {
        uint64_t p = base + lane;
        int alive = 0;
        uint64_t cur = p;
        uint32_t cnt = 0;
        int complete = 1;
        if (p < T) {
            alive = 1;
            for (;;) {
                Step s = chain_cand<F>(a.buf, len, cur, T, a.samples, a.contigs);
                if (s.kind == 1) break;
                if (s.kind == 3) { alive = 0; break; }
                ++cnt;
                cur = s.next;
            }
            if (F == CHAIN_BAM && alive && !landing_plausible(a.buf, len, cur)) alive = 0;
        }
        uint64_t m = __ballot(alive);
        if (m) {
            uint32_t first = __ffsll((unsigned long long)m) - 1;
            if (n_surv == 0) {
                L = uni64(__shfl(cur, first, 64));
                Lcomplete = (int)uni((uint32_t)__shfl(complete, first, 64));
            }
            if (alive && (cur != L || complete != Lcomplete)) agree = 0;
            uint32_t rank = n_surv + (uint32_t)__popcll(m & lt);
            if (alive && rank < BAM_K) {
                a.surv_pos[(uint64_t)b * BAM_K + rank] = p;
                a.surv_cnt[(uint64_t)b * BAM_K + rank] = cnt;
            }
            n_surv += (uint32_t)__popcll(m);
        }
        if (base + 64 >= T && n_surv == 0 && T < sb1) {
            // no survivor yet: the next window starts at the first position
            // that could begin a record (BAM), or right after this one (BCF)
            const uint64_t q = F == CHAIN_BAM ? skim_bam(a.buf, len, T, sb1, lane) : T;
            if (q >= sb1) {
                T = sb1;
                base = sb1;
            } else {
                base = q - 64; // the loop step moves it to q
                T = (sb1 - q) > a.window ? q + a.window : sb1;
            }
        }
    }
    agree = (int)uni((uint32_t)(__ballot(!agree) == 0));
    if (n_surv == 0 || !agree) {
        if (lane == 0) {
            a.spec_ok[b] = 0;
            a.n_surv[b] = (uint8_t)(n_surv < BAM_K ? n_surv : BAM_K);
        }
        return;
    }
    // the exact walk from the agreed landing is bam_walk_kernel's
    if (lane == 0) {
        a.spec_ok[b] = 1;
        a.n_surv[b] = (uint8_t)(n_surv < BAM_K ? n_surv : BAM_K);
        a.land[b] = L;
        a.complete[b] = (uint8_t)Lcomplete;
    }
}

// The exact walk of every member whose speculation agreed (spec_ok 1 from
// bam_spec_kernel), from its landing to its end, one lane per member: the
// walk is a chain of dependent loads, and a wave keeps 64 of them in flight.
template <uint32_t F>
__global__ __launch_bounds__(64) void bam_walk_kernel(BamSpecArgs a)
{
    const uint32_t b = blockIdx.x * 64 + threadIdx.x;
    if (b >= a.nb || a.spec_ok[b] != 1) return;
    const uint64_t sb = a.block_starts[b];
    const uint64_t sb1 = a.mem_until ? a.mem_until[b] : a.block_starts[b + 1];
    const uint64_t len = a.mem_len ? a.mem_len[b] : a.len;
    uint64_t cur = a.land[b];
    uint32_t cnt = 0;
    int complete = a.complete[b], valid = 1;
    uint16_t *rel = a.rel_offs + (uint64_t)b * BAM_CAPB;
    if (complete) {
        // The offsets go out eight at a time as one 16-byte store (the slab
        // is 16-byte aligned): a store in every step would make every step's
        // load wait for it too (the memory counter covers both). q0:q1 is a
        // shift register of eight u16, the newest at the top.
        uint64_t q0 = 0, q1 = 0;
        for (;;) {
            Step s = chain_step<F>(a.buf, len, cur, sb1);
            if (s.kind == 1) break;
            if (s.kind == 2) { complete = 0; break; }
            if (s.kind == 3) { valid = 0; break; }
            q0 = (q0 >> 16) | (q1 << 48);
            q1 = (q1 >> 16) | ((uint64_t)(uint16_t)(cur - sb) << 48);
            ++cnt;
            if ((cnt & 7) == 0)
                *reinterpret_cast<uint4 *>(rel + cnt - 8) =
                    make_uint4((uint32_t)q0, (uint32_t)(q0 >> 32), (uint32_t)q1, (uint32_t)(q1 >> 32));
            cur = s.next;
        }
        // the last cnt % 8 offsets sit in the top slots of the register
        const uint32_t k = cnt & 7;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t slot = 8 - k + j;
            rel[cnt - k + j] = (uint16_t)(slot < 4 ? q0 >> (16 * slot) : q1 >> (16 * (slot - 4)));
        }
    }
    a.spec_ok[b] = (uint8_t)valid;
    a.exitp[b] = cur;
    a.cnt_tail[b] = cnt;
    a.complete[b] = (uint8_t)complete;
}

// Lane j of the result takes v, the other lanes keep old (v is wave-uniform).
__device__ __forceinline__ uint32_t wl32(uint32_t v, uint32_t j, uint32_t old)
{
    return threadIdx.x == j ? v : old;
}
__device__ __forceinline__ uint64_t wl64(uint64_t v, uint32_t j, uint64_t old)
{
    return threadIdx.x == j ? v : old;
}
__device__ __forceinline__ uint64_t rl64(uint64_t v, uint32_t j)
{
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)j);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)j);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t rl32(uint32_t v, uint32_t j)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}

// The true chain over members, 64 members per step (one wave). Each lane first
// checks, in parallel, whether its member can be adopted assuming the previous
// member was; the uniform pass then follows the chain with register reads and
// walks only the members whose speculation does not apply.
template <uint32_t F>
__global__ __launch_bounds__(64) void bam_reconcile_kernel(BamReconcileArgs a)
{
    if (!a.seg_mb && a.fast_state && a.fast_state[2]) return; // the parallel path did it
    // (the parallel path runs only without mem_until, where s1 is the next block start)
    const uint32_t lane = threadIdx.x;
    uint32_t mb0 = 0, mb1 = a.nb;
    uint64_t *totals = a.totals;
    uint64_t cur;
    if (a.seg_mb) {
        const uint32_t q = blockIdx.x;
        mb0 = a.seg_mb[q];
        mb1 = a.seg_mb[q + 1];
        a.len = a.seg_len[q];
        cur = a.seg_start[q];
        totals = a.seg_totals + 4ull * q;
    } else {
        if (a.len_dev) a.len = *a.len_dev;
        cur = a.start_dev ? *a.start_dev : a.start;
    }
    uint64_t emitted = 0, tail = a.len, fallbacks = 0;
    int stopped = 0;
    uint64_t status = BAM_OK;
    for (uint32_t base = mb0; base < mb1; base += 64) {
        const uint32_t b = base + lane;
        const bool in = b < mb1;
        uint64_t s1 = in ? (a.mem_until ? a.mem_until[b] : a.block_starts[b + 1]) : 0;
        uint32_t ok = in ? a.spec_ok[b] : 0;
        uint64_t X = in && ok ? a.exitp[b] : 0;
        uint32_t comp = in && ok ? a.complete[b] : 0;
        uint32_t ctail = in && ok ? a.cnt_tail[b] : 0;
        // expected entry = previous lane's speculative exit (lane 0: the chain)
        uint64_t prevX = __shfl_up(X, 1, 64);
        uint64_t expect = lane == 0 ? cur : prevX;
        uint32_t match = 0, mcnt = 0;
        if (ok) {
            uint32_t ns = a.n_surv[b];
            for (uint32_t k = 0; k < ns; ++k) {
                if (a.surv_pos[(uint64_t)b * BAM_K + k] == expect) {
                    match = k + 1;
                    mcnt = a.surv_cnt[(uint64_t)b * BAM_K + k] + ctail;
                    break;
                }
            }
        }
        uint64_t o_entry = BAM_NONE, o_first = 0;
        uint32_t o_cnt = 0, o_mode = 0;
        uint32_t lim = mb1 - base < 64 ? mb1 - base : 64;
        for (uint32_t j = 0; j < lim; ++j) {
            uint64_t js1 = rl64(s1, j);
            if (stopped || cur >= js1) {
                o_first = wl64(emitted, j, o_first);
                continue;
            }
            uint32_t cnt, comp_j, valid = 1, mode;
            uint64_t land;
            uint32_t found = 0, fcnt = 0;
            if (rl32(match, j) && rl64(expect, j) == cur) {
                found = rl32(match, j);
                fcnt = rl32(mcnt, j);
            } else if (rl32(ok, j)) {
                // the chain did not arrive where the previous member's
                // speculation said; look for it among this member's survivors
                const uint32_t bj = base + j;
                const uint32_t ns = a.n_surv[bj];
                for (uint32_t k = 0; k < ns && !found; ++k) {
                    if (a.surv_pos[(uint64_t)bj * BAM_K + k] == cur) {
                        found = k + 1;
                        fcnt = a.surv_cnt[(uint64_t)bj * BAM_K + k] + rl32(ctail, j);
                    }
                }
            }
            if (found) {
                cnt = fcnt;
                land = rl64(X, j);
                comp_j = rl32(comp, j);
                mode = found;
            } else {
                Walk w = walk_uniform<F>(a.buf, a.len, cur, js1);
                cnt = w.count;
                land = w.land;
                comp_j = (uint32_t)w.complete;
                valid = (uint32_t)w.valid;
                mode = BAM_MODE_FALLBACK;
                ++fallbacks;
            }
            o_first = wl64(emitted, j, o_first);
            if (!valid) {
                status = BAM_MALFORMED;
                // BAM reports the chain's entry into the member (blocked.rs);
                // BCF the record itself (record.rs scan_records)
                tail = F == CHAIN_BCF ? land : cur;
                stopped = 1;
                continue;
            }
            o_entry = wl64(cur, j, o_entry);
            o_cnt = wl32(cnt, j, o_cnt);
            o_mode = wl32(mode, j, o_mode);
            emitted += cnt;
            cur = land;
            if (!comp_j) {
                tail = land;
                stopped = 1;
            }
        }
        if (in) {
            a.entry[b] = o_entry;
            a.count[b] = o_cnt;
            a.first_index[b] = o_first;
            a.mode[b] = (uint8_t)o_mode;
        }
    }
    if (!stopped) tail = cur;
    if (lane == 0) {
        totals[0] = emitted;
        totals[1] = tail;
        totals[2] = status;
        totals[3] = fallbacks;
    }
}

// Parallel reconcile for one chain, in the common case where every member is
// entered at the previous member's speculative exit and that position is one
// of its survivors: each member's count and mode follow from its own
// speculation, first_index from a scan of the counts. The check kernel flags
// the first member where the assumption fails (state[0]) and the first where
// the chain ends inside the member (state[1]); the finish kernel writes the
// outputs when no failure comes before the chain's end and sets state[2], and
// the serial reconcile that runs next returns at once. Otherwise the serial
// reconcile produces every output, as without this path.
__global__ void bam_fast_init_kernel(uint32_t *st)
{
    if (threadIdx.x == 0) {
        st[0] = ~0u;
        st[1] = ~0u;
        st[2] = 0;
    }
}

__global__ __launch_bounds__(256) void bam_fast_check_kernel(BamReconcileArgs a)
{
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j >= a.nb) return;
    const uint64_t s1 = a.block_starts[j + 1];
    if (s1 == a.block_starts[j]) {
        // an empty member (the EOF block): the chain passes it by
        a.entry[j] = BAM_NONE;
        a.count[j] = 0;
        a.mode[j] = 0;
        return;
    }
    // entered at the speculative exit of the closest earlier member holding a
    // record start; the members between hold none (their speculation scanned
    // them whole and found no candidate), so the chain passes them by, which
    // each of them checks for itself below
    uint64_t e = a.start_dev ? *a.start_dev : a.start;
    bool ok = true;
    for (uint32_t i = j, hops = 0; i > 0; --i) {
        const uint32_t k = i - 1;
        if (a.block_starts[k + 1] == a.block_starts[k]) continue;
        if (a.n_surv[k] == 0) {
            if (++hops > FAST_MAX_HOPS) {
                ok = false;
                break;
            }
            continue;
        }
        ok = a.spec_ok[k] != 0;
        e = a.exitp[k];
        break;
    }
    if (a.n_surv[j] == 0) {
        // no record starts here: the chain must pass the member by
        a.entry[j] = BAM_NONE;
        a.count[j] = 0;
        a.mode[j] = 0;
        if (!ok || e < s1) atomicMin(&a.fast_state[0], j);
        return;
    }
    ok = ok && a.spec_ok[j] != 0;
    uint32_t found = 0, cnt = 0;
    if (ok && e < s1) {
        const uint32_t ns = a.n_surv[j];
        for (uint32_t k = 0; k < ns && !found; ++k) {
            if (a.surv_pos[(uint64_t)j * BAM_K + k] == e) {
                found = k + 1;
                cnt = a.surv_cnt[(uint64_t)j * BAM_K + k] + a.cnt_tail[j];
            }
        }
    }
    a.entry[j] = e;
    a.count[j] = cnt;
    a.mode[j] = (uint8_t)found;
    if (!found)
        atomicMin(&a.fast_state[0], j);
    else if (!a.complete[j])
        atomicMin(&a.fast_state[1], j);
}

__global__ __launch_bounds__(256) void bam_fast_finish_kernel(BamReconcileArgs a)
{
    const uint32_t fail = a.fast_state[0], stop = a.fast_state[1];
    const uint32_t last = stop == ~0u ? a.nb - 1 : stop;
    if (fail != ~0u && fail <= last) return;
    if (a.block_starts[a.nb] == a.block_starts[0]) return; // no member holds data
    const uint32_t j = blockIdx.x * 256 + threadIdx.x;
    if (j < a.nb) {
        if (j <= last) {
            a.first_index[j] = a.fast_scan[j];
        } else {
            a.first_index[j] = a.fast_scan[last + 1];
            a.entry[j] = BAM_NONE;
            a.count[j] = 0;
            a.mode[j] = 0;
        }
    }
    if (j == 0) {
        // tail: the exit of the last member the chain entered (members it
        // passed by, and empty ones, have no exit of their own)
        uint32_t t = last;
        while (t > 0 && a.entry[t] == BAM_NONE) --t;
        a.totals[0] = a.fast_scan[last + 1];
        a.totals[1] = a.entry[t] == BAM_NONE ? (a.start_dev ? *a.start_dev : a.start) : a.exitp[t];
        a.totals[2] = BAM_OK;
        a.totals[3] = 0;
        a.fast_state[2] = 1;
    }
}

__global__ void bam_segment_bases_kernel(const uint64_t *seg_totals, uint32_t n_seg, uint64_t *seg_base)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t run = 0;
    for (uint32_t q = 0; q < n_seg; ++q) {
        seg_base[q] = run;
        run += seg_totals[4ull * q];
    }
    seg_base[n_seg] = run;
}

template <uint32_t F>
__global__ __launch_bounds__(64) void bam_emit_kernel(BamEmitArgs a)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t b = blockIdx.x;
    const uint64_t e = a.entry[b];
    if (e == BAM_NONE) return;
    const uint32_t cnt = a.count[b];
    const uint64_t first = a.first_index[b] + (a.seg_base ? a.seg_base[a.mem_seg[b]] : 0);
    const uint32_t mode = a.mode[b];
    const uint64_t sb = a.block_starts[b];
    const uint64_t vbase = a.coffsets ? (a.coffsets[b] << 16) : 0;
    uint32_t nwalk = cnt;
    if (mode != BAM_MODE_FALLBACK) nwalk = a.surv_cnt[(uint64_t)b * BAM_K + mode - 1];
    uint64_t cur = e;
    for (uint32_t i = 0; i < nwalk; ++i) {
        if (lane == 0) {
            a.record_offsets[first + i] = cur;
            if (a.virtual_offsets) a.virtual_offsets[first + i] = vbase | (cur - sb);
        }
        cur = chain_next<F>(a.buf, cur);
    }
    if (mode != BAM_MODE_FALLBACK) {
        const uint16_t *rel = a.rel_offs + (uint64_t)b * BAM_CAPB;
        for (uint32_t i = lane; i < cnt - nwalk; i += 64) {
            uint64_t r = rel[i];
            a.record_offsets[first + nwalk + i] = sb + r;
            if (a.virtual_offsets) a.virtual_offsets[first + nwalk + i] = vbase | r;
        }
    }
}

__global__ __launch_bounds__(256) void bam_decode_kernel(BamDecodeArgs a)
{
    const uint64_t n = a.n_dev ? *a.n_dev : a.n;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t r = a.c.record_offsets[i];
        uint32_t w[9];
        fixed_words(a.buf, a.len, r, w);
        const uint32_t bs = w[0];
        const uint32_t name_len = w[3] & 0xFF;
        const uint32_t nc = w[4] & 0xFFFF;
        const uint32_t ls = w[5];
        a.c.ref_id[i] = (int32_t)w[1];
        a.c.pos[i] = (int32_t)w[2];
        a.c.mapq[i] = (uint8_t)(w[3] >> 8);
        a.c.flags[i] = (uint16_t)(w[4] >> 16);
        a.c.l_seq[i] = ls;
        a.c.next_ref_id[i] = (int32_t)w[6];
        a.c.next_pos[i] = (int32_t)w[7];
        a.c.tlen[i] = (int32_t)w[8];
        uint32_t c = 36u + name_len;
        uint32_t s = c + 4u * nc;
        uint32_t q = s + (ls + 1u) / 2u;
        a.c.cigar_start[i] = c;
        a.c.seq_start[i] = s;
        a.c.qual_start[i] = q;
        a.c.aux_start[i] = q + ls;
        a.c.record_end[i] = 4u + bs;
    }
}

// BCF typed value (crates/fritillaria-bcf/src/columnar/typed.rs:255-300): bytes
// per element of a type code, or 0xFFFFFFFF for a reserved code.
__device__ __forceinline__ uint32_t bcf_type_size(uint32_t code)
{
    switch (code) {
    case 0: return 0;
    case 1: case 7: return 1;
    case 2: return 2;
    case 3: case 5: return 4;
    default: return 0xFFFFFFFFu;
    }
}

// End of the typed value at pos inside [.., end), or 0 when it does not parse:
// a reserved type, an escaped count that is not an atomic integer >= 15, or a
// value running past end.
__device__ __forceinline__ uint64_t bcf_skip(const uint8_t *buf, uint64_t end, uint64_t pos)
{
    if (pos >= end) return 0;
    const uint32_t d = buf[pos], code = d & 15;
    const uint32_t size = bcf_type_size(code);
    if (size == 0xFFFFFFFFu) return 0;
    uint64_t count = d >> 4, header = 1;
    if (count == 15) {
        if (pos + 1 >= end) return 0;
        const uint32_t in = buf[pos + 1], ic = in & 15;
        if ((in >> 4) != 1 || (ic != 1 && ic != 2 && ic != 3)) return 0;
        const uint32_t w = bcf_type_size(ic);
        if (pos + 2 + w > end) return 0;
        const uint8_t *p = buf + pos + 2;
        const int32_t v = ic == 1 ? (int32_t)(int8_t)p[0] : ic == 2 ? (int32_t)(int16_t)ld16(p) : (int32_t)ld32(p);
        if (v < 15) return 0;
        count = (uint64_t)v;
        header = 2 + w;
    }
    const uint64_t e = pos + header + count * size;
    return e > end ? 0 : e;
}

// RecordBatch::decode for BCF (crates/fritillaria-bcf/src/columnar/batch.rs:158-172,
// record.rs Record::bounds): the site core, then ID, the alleles and FILTER
// skipped as typed values inside the record. A record whose values do not
// parse is reported by index (the lowest one wins).
__global__ __launch_bounds__(256) void bcf_decode_kernel(BcfDecodeArgs a)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t o = a.c.record_offsets[i];
        const uint64_t e = i + 1 < a.n ? a.c.record_offsets[i + 1] : a.tail;
        const uint8_t *r = a.buf + o;
        const uint32_t ls = ld32(r), li = ld32(r + 4);
        const uint32_t n_allele = ld16(r + 26);
        a.c.chrom[i] = (int32_t)ld32(r + 8);
        a.c.pos[i] = (int32_t)ld32(r + 12);
        a.c.rlen[i] = (int32_t)ld32(r + 16);
        a.c.qual[i] = ld32(r + 20);
        a.c.n_info[i] = (uint16_t)ld16(r + 24);
        a.c.n_allele[i] = (uint16_t)n_allele;
        a.c.n_sample[i] = ld16(r + 28) | ((uint32_t)r[30] << 16);
        a.c.n_fmt[i] = r[31];
        a.c.genotypes_start[i] = 8 + ls;
        a.c.record_end[i] = (uint32_t)(e - o);
        bool ok = e - o >= 32 && 8ull + ls + li == e - o;
        uint64_t at = ok ? bcf_skip(a.buf, e, o + 32) : 0;
        ok = ok && at != 0;
        a.c.alleles_start[i] = ok ? (uint32_t)(at - o) : 0;
        for (uint32_t k = 0; ok && k < n_allele; ++k) {
            at = bcf_skip(a.buf, e, at);
            ok = at != 0;
        }
        a.c.filters_start[i] = ok ? (uint32_t)(at - o) : 0;
        if (ok) {
            at = bcf_skip(a.buf, e, at);
            ok = at != 0;
        }
        a.c.info_start[i] = ok ? (uint32_t)(at - o) : 0;
        if (!ok) atomicMin(a.first_bad, (unsigned long long)i);
    }
}

} // namespace

hipError_t launch_bcf_decode(const BcfDecodeArgs &a, hipStream_t s)
{
    if (a.n == 0) return hipSuccess;
    uint64_t blocks = (a.n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(bcf_decode_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bam_spec(const BamSpecArgs &a, hipStream_t s)
{
    if (a.nb == 0) return hipSuccess;
    // the walk writes each member's offsets 16 bytes at a time into its slab
    if (((uintptr_t)a.rel_offs & 15) != 0) return hipErrorInvalidValue;
    // one-wave workgroups: the walks are latency-bound chains, spread over every CU
    const uint32_t wgrid = (a.nb + 63) / 64;
    if (a.format == CHAIN_BCF) {
        hipLaunchKernelGGL(bam_spec_kernel<CHAIN_BCF>, dim3(a.nb), dim3(64), 0, s, a);
        hipLaunchKernelGGL(bam_walk_kernel<CHAIN_BCF>, dim3(wgrid), dim3(64), 0, s, a);
    } else {
        hipLaunchKernelGGL(bam_spec_kernel<CHAIN_BAM>, dim3(a.nb), dim3(64), 0, s, a);
        hipLaunchKernelGGL(bam_walk_kernel<CHAIN_BAM>, dim3(wgrid), dim3(64), 0, s, a);
    }
    return hipGetLastError();
}

hipError_t launch_bam_reconcile(const BamReconcileArgs &a, hipStream_t s)
{
    if (a.fast_state && a.fast_scan && a.nb > 0 && !a.mem_until) {
        const uint32_t grid = (a.nb + 255) / 256;
        hipLaunchKernelGGL(bam_fast_init_kernel, dim3(1), dim3(64), 0, s, a.fast_state);
        hipLaunchKernelGGL(bam_fast_check_kernel, dim3(grid), dim3(256), 0, s, a);
        hipError_t r = hipGetLastError();
        if (r == hipSuccess) r = launch_scan_u32_u64(a.count, a.fast_scan, nullptr, a.nb, s);
        if (r != hipSuccess) return r;
        hipLaunchKernelGGL(bam_fast_finish_kernel, dim3(grid), dim3(256), 0, s, a);
        r = hipGetLastError();
        if (r != hipSuccess) return r;
    }
    if (a.format == CHAIN_BCF)
        hipLaunchKernelGGL(bam_reconcile_kernel<CHAIN_BCF>, dim3(1), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(bam_reconcile_kernel<CHAIN_BAM>, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bam_reconcile_segments(const BamReconcileArgs &a, uint32_t n_seg, hipStream_t s)
{
    if (n_seg == 0) return hipSuccess;
    if (a.format == CHAIN_BCF)
        hipLaunchKernelGGL(bam_reconcile_kernel<CHAIN_BCF>, dim3(n_seg), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(bam_reconcile_kernel<CHAIN_BAM>, dim3(n_seg), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bam_segment_bases(const uint64_t *seg_totals, uint32_t n_seg, uint64_t *seg_base, hipStream_t s)
{
    hipLaunchKernelGGL(bam_segment_bases_kernel, dim3(1), dim3(64), 0, s, seg_totals, n_seg, seg_base);
    return hipGetLastError();
}

hipError_t launch_bam_emit(const BamEmitArgs &a, hipStream_t s)
{
    if (a.nb == 0) return hipSuccess;
    if (a.format == CHAIN_BCF)
        hipLaunchKernelGGL(bam_emit_kernel<CHAIN_BCF>, dim3(a.nb), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL(bam_emit_kernel<CHAIN_BAM>, dim3(a.nb), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_bam_decode(const BamDecodeArgs &a, uint64_t max_records, hipStream_t s)
{
    if (max_records == 0) return hipSuccess;
    uint64_t blocks = (max_records + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(bam_decode_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
