// fastq.hip -- FASTQ record boundaries and line-bound columns, on the device.
//
// Semantics are the serial walk of crates/fritillaria-fastq/src/columnar/record.rs:145-176
// (newlines between records are skipped; a position that is not a whole record
// ends the walk, as the tail when it holds '@', as a Malformed error otherwise)
// with the record test of record.rs:85-131. The device path of the reference
// (crates/fritillaria-cuda/src/fastq.rs:469-535, kernels/fastq_scan.cu) sieves
// line starts, sorts them on the host, proves a tiling and falls back to a
// one-thread walk. Here the walk itself is computed in parallel and exactly:
//
//   1. newline index: per 64 KiB tile a count, a scan, an ordered write of
//      every '\n' position at or after start (one read of the batch);
//   2. line types: one thread per line classifies it as BLANK (skipped),
//      RECORD (the four lines from here form a record: the walk jumps 4
//      lines) or STOP (the walk ends here); lines past the end are END;
//   3. the walk moves +1 or +4 lines, so entering a segment of 256 lines it
//      can only be at one of 4 entry lines. Per segment, 4 walkers (one per
//      entry) give the exit entry and the record count: a function on 4
//      states;
//   4. one workgroup composes those functions (chunked, then a Hillis-Steele
//      scan over 1024 chunk functions), giving every segment its true entry
//      and record index base;
//   5. each segment walks once more from its true entry and writes the
//      record offsets and line bounds; the segment holding the end of the
//      walk writes the tail and the error status.
//
// No speculation, so no fallback: the result is the serial walk's on every
// input, blank lines and decoys included.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fastq_internal.h"

namespace frhip {
namespace {

__device__ __forceinline__ uint32_t nl_mask16(uint4 v)
{
    // bit i set when byte i of the 16 is '\n' (exact per byte, no borrow carry)
    uint32_t m = 0;
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint32_t x = w[q] ^ 0x0A0A0A0Au;
        uint32_t t = ((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x | 0x7F7F7F7Fu;
        t = ~t; // 0x80 in each byte that was '\n'
        m |= (((t >> 7) & 1u) | ((t >> 14) & 2u) | ((t >> 21) & 4u) | ((t >> 28) & 8u)) << (4 * q);
    }
    return m;
}

// Newline mask of the 16 bytes at absolute offset c (16-aligned), restricted
// to [start, len). Bytes at or past len are never read.
__device__ __forceinline__ uint32_t chunk_mask(const uint8_t *buf, uint64_t start, uint64_t len, uint64_t c)
{
    if (c >= len || c + 16 <= start) return 0;
    uint32_t m;
    if (c + 16 <= len) {
        m = nl_mask16(*reinterpret_cast<const uint4 *>(buf + c));
    } else {
        m = 0;
        for (uint32_t i = 0; c + i < len; ++i) m |= (buf[c + i] == '\n' ? 1u : 0u) << i;
    }
    if (c < start) m &= ~0u << (uint32_t)(start - c);
    return m & 0xFFFFu;
}

__global__ __launch_bounds__(256) void fq_count_kernel(FastqArgs a)
{
    __shared__ uint32_t part[4];
    const uint64_t tile_lo = (a.first_tile + blockIdx.x) * (uint64_t)FQ_TILE;
    uint32_t n = 0;
    for (uint32_t j = 0; j < FQ_TILE / 4096; ++j)
        n += __popc(chunk_mask(a.buf, a.start, a.len, tile_lo + j * 4096u + threadIdx.x * 16u));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = n;
    __syncthreads();
    if (threadIdx.x == 0) a.tile_counts[blockIdx.x] = part[0] + part[1] + part[2] + part[3];
}

__global__ __launch_bounds__(256) void fq_write_kernel(FastqArgs a)
{
    __shared__ uint32_t wsum[4];
    const uint64_t tile_lo = (a.first_tile + blockIdx.x) * (uint64_t)FQ_TILE;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint64_t base = a.tile_base[blockIdx.x];
    for (uint32_t j = 0; j < FQ_TILE / 4096; ++j) {
        const uint64_t c = tile_lo + j * 4096u + threadIdx.x * 16u;
        uint32_t m = chunk_mask(a.buf, a.start, a.len, c);
        uint32_t cnt = __popc(m);
        // inclusive scan within the wave
        uint32_t inc = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            uint32_t y = __shfl_up(inc, o, 64);
            if (lane >= (uint32_t)o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (uint32_t w = 0; w < 4; ++w) {
            before += w < wave ? wsum[w] : 0u;
            total += wsum[w];
        }
        uint64_t o = base + before + inc - cnt;
        while (m) {
            uint32_t b = __ffs(m) - 1;
            m &= m - 1;
            a.nl[o++] = c + b;
        }
        base += total;
        __syncthreads();
    }
}

__device__ __forceinline__ uint64_t line_start(const FastqArgs &a, uint64_t k)
{
    if (k == 0) return a.start;
    if (k > a.n_nl) return a.len;
    return a.nl[k - 1] + 1;
}

__global__ __launch_bounds__(256) void fq_type_kernel(FastqArgs a)
{
    __shared__ uint32_t part[4];
    uint32_t surv = 0;
    for (uint64_t k = blockIdx.x * 256ull + threadIdx.x; k < a.n_lines_pad; k += (uint64_t)gridDim.x * 256) {
        uint8_t t;
        if (k >= a.n_lines) {
            t = FQ_END;
        } else {
            const uint64_t lk = line_start(a, k);
            if (k < a.n_nl && a.nl[k] == lk) {
                t = FQ_BLANK;
            } else if (k + 2 < a.n_nl && a.buf[lk] == '@' && a.buf[a.nl[k + 1] + 1] == '+') {
                const uint64_t seq_len = a.nl[k + 1] - a.nl[k] - 1;
                const uint64_t q0 = a.nl[k + 2] + 1;
                const uint64_t q1 = k + 3 < a.n_nl ? a.nl[k + 3] : a.len;
                t = q1 - q0 == seq_len ? FQ_RECORD : FQ_STOP;
            } else {
                t = FQ_STOP;
            }
        }
        surv += t == FQ_RECORD;
        a.type[k] = t;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) surv += __shfl_xor(surv, o, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = surv;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = part[0] + part[1] + part[2] + part[3];
        if (s) atomicAdd(reinterpret_cast<unsigned long long *>(&a.totals[FQ_T_SURVIVORS]), (unsigned long long)s);
    }
}

// Per segment, the walk from each of the 4 entry lines: exit entry (0-3) or
// FQ_EXIT_DONE, and the records met. Packed 16 bits per entry:
// exit in bits 0-2, count in bits 3-15.
__global__ __launch_bounds__(256) void fq_segment_kernel(FastqArgs a)
{
    __shared__ uint8_t ty[64 * FQ_SEG];
    const uint64_t seg0 = blockIdx.x * 64ull;
    const uint64_t line0 = seg0 * FQ_SEG;
    // 16 KiB of types, n_lines_pad is a multiple of it
    const uint4 *src = reinterpret_cast<const uint4 *>(a.type + line0);
    uint4 *dst = reinterpret_cast<uint4 *>(ty);
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[threadIdx.x + 256 * i] = src[threadIdx.x + 256 * i];
    __syncthreads();
    const uint32_t s = threadIdx.x >> 2, e = threadIdx.x & 3;
    const uint8_t *t = ty + s * FQ_SEG;
    uint32_t k = e, cnt = 0, exitv = FQ_EXIT_DONE;
    while (k < FQ_SEG) {
        const uint8_t v = t[k];
        if (v == FQ_BLANK) {
            k += 1;
        } else if (v == FQ_RECORD) {
            ++cnt;
            k += 4;
        } else {
            break;
        }
    }
    if (k >= FQ_SEG) exitv = k - FQ_SEG;
    const uint32_t packed = exitv | (cnt << 3);
    // gather the 4 entries of the segment into one 64-bit word
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t g = lane & ~3u;
    const uint64_t w = (uint64_t)(__shfl(packed, g, 64) & 0xFFFF) | ((uint64_t)(__shfl(packed, g + 1, 64) & 0xFFFF) << 16) |
                       ((uint64_t)(__shfl(packed, g + 2, 64) & 0xFFFF) << 32) |
                       ((uint64_t)(__shfl(packed, g + 3, 64) & 0xFFFF) << 48);
    if (e == 0 && seg0 + s < a.n_seg) a.seg_f[seg0 + s] = w;
}

struct Fn {
    uint8_t exit[4];
    uint32_t cnt[4];
};

__device__ __forceinline__ Fn fn_identity()
{
    Fn f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        f.exit[e] = (uint8_t)e;
        f.cnt[e] = 0;
    }
    return f;
}

// then(f, g): apply f, then g.
__device__ __forceinline__ Fn fn_then(const Fn &f, const Fn &g)
{
    Fn r;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t x = f.exit[e];
        if (x < 4) {
            r.exit[e] = g.exit[x];
            r.cnt[e] = f.cnt[e] + g.cnt[x];
        } else {
            r.exit[e] = (uint8_t)x;
            r.cnt[e] = f.cnt[e];
        }
    }
    return r;
}

__device__ __forceinline__ Fn fn_unpack(uint64_t w)
{
    Fn f;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint32_t p = (uint32_t)(w >> (16 * e)) & 0xFFFFu;
        f.exit[e] = (uint8_t)(p & 7u);
        f.cnt[e] = p >> 3;
    }
    return f;
}

__global__ __launch_bounds__(1024) void fq_compose_kernel(FastqArgs a)
{
    __shared__ uint8_t sx[2][1024][4];
    __shared__ uint32_t sc[2][1024][4];
    const uint32_t t = threadIdx.x;
    const uint64_t ch = (a.n_seg + 1023) / 1024;
    const uint64_t s0 = t * ch, s1 = s0 + ch < a.n_seg ? s0 + ch : a.n_seg;
    Fn f = fn_identity();
    for (uint64_t s = s0; s < s1; ++s) f = fn_then(f, fn_unpack(a.seg_f[s]));
    int cur = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        sx[0][t][e] = f.exit[e];
        sc[0][t][e] = f.cnt[e];
    }
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        Fn mine;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            mine.exit[e] = sx[cur][t][e];
            mine.cnt[e] = sc[cur][t][e];
        }
        if (t >= d) {
            Fn prev;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                prev.exit[e] = sx[cur][t - d][e];
                prev.cnt[e] = sc[cur][t - d][e];
            }
            mine = fn_then(prev, mine);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            sx[cur ^ 1][t][e] = mine.exit[e];
            sc[cur ^ 1][t][e] = mine.cnt[e];
        }
        cur ^= 1;
        __syncthreads();
    }
    // state entering this chunk: the inclusive prefix of the chunks before, at entry 0
    uint32_t state = 0;
    uint64_t base = 0;
    if (t > 0) {
        state = sx[cur][t - 1][0];
        base = sc[cur][t - 1][0];
    }
    for (uint64_t s = s0; s < s1; ++s) {
        a.seg_state[s] = (uint8_t)state;
        a.seg_base[s] = base;
        if (state < 4) {
            const uint64_t w = a.seg_f[s];
            const uint32_t p = (uint32_t)(w >> (16 * state)) & 0xFFFFu;
            base += p >> 3;
            state = p & 7u;
        }
    }
    if (s0 < s1 && s1 == a.n_seg) a.totals[FQ_T_RECORDS] = base;
}

__global__ __launch_bounds__(64) void fq_emit_kernel(FastqArgs a)
{
    const uint64_t s = blockIdx.x * 64ull + threadIdx.x;
    if (s >= a.n_seg) return;
    const uint32_t state = a.seg_state[s];
    if (state >= 4) return;
    uint64_t idx = a.seg_base[s];
    const uint64_t seg_lo = s * FQ_SEG, seg_hi = seg_lo + FQ_SEG;
    uint64_t k = seg_lo + state;
    while (k < seg_hi) {
        const uint8_t v = a.type[k];
        if (v == FQ_BLANK) {
            k += 1;
        } else if (v == FQ_RECORD) {
            if (idx < a.cap) {
                const uint64_t lk = line_start(a, k);
                a.record_offsets[idx] = lk;
                a.sequence_start[idx] = (uint32_t)(a.nl[k] + 1 - lk);
                a.plus_start[idx] = (uint32_t)(a.nl[k + 1] + 1 - lk);
                a.quality_start[idx] = (uint32_t)(a.nl[k + 2] + 1 - lk);
                a.sequence_len[idx] = (uint32_t)(a.nl[k + 1] - a.nl[k] - 1);
                const uint64_t end = k + 3 < a.n_nl ? a.nl[k + 3] + 1 : a.len;
                a.record_end[idx] = (uint32_t)(end - lk);
            }
            ++idx;
            k += 4;
        } else {
            const uint64_t lk = line_start(a, k);
            a.totals[FQ_T_TAIL] = lk;
            a.totals[FQ_T_STATUS] = (v == FQ_STOP && a.buf[lk] != '@') ? 1 : 0;
            return;
        }
    }
}

} // namespace

uint64_t fastq_tiles(uint64_t start, uint64_t len)
{
    if (start >= len) return 0;
    return (len - 1) / FQ_TILE - start / FQ_TILE + 1;
}

hipError_t launch_fastq_index(const FastqArgs &a, hipStream_t s)
{
    const uint32_t nt = (uint32_t)a.n_tiles;
    hipLaunchKernelGGL(fq_count_kernel, dim3(nt), dim3(256), 0, s, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    return launch_scan_u32_u64(a.tile_counts, a.tile_base, nullptr, a.n_tiles, s);
}

hipError_t launch_fastq_lines(const FastqArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(fq_write_kernel, dim3((uint32_t)a.n_tiles), dim3(256), 0, s, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    uint64_t g = (a.n_lines_pad + 255) / 256;
    if (g > 8192) g = 8192;
    hipLaunchKernelGGL(fq_type_kernel, dim3((uint32_t)g), dim3(256), 0, s, a);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    hipLaunchKernelGGL(fq_segment_kernel, dim3((uint32_t)(a.n_lines_pad / (64 * FQ_SEG))), dim3(256), 0, s, a);
    if ((r = hipGetLastError()) != hipSuccess) return r;
    hipLaunchKernelGGL(fq_compose_kernel, dim3(1), dim3(1024), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fastq_emit(const FastqArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(fq_emit_kernel, dim3((uint32_t)((a.n_seg + 63) / 64)), dim3(64), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
