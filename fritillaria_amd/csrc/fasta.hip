// fasta.hip -- a FASTA reference made contiguous in device memory.
//
// FastaCompactor::compact (crates/fritillaria-cuda/src/fasta.rs) finds contig
// starts ('>' at a line start) on the device, then copies the whole file back
// to the host to lay the contigs out (crates/fritillaria-fasta/src/columnar/
// record.rs:58-126: a contig's sequence runs from after its definition line to
// the next contig start, its bases are the bytes that are not newlines; a last
// '>' whose definition line has no newline is not a record), and compacts each
// contig with a per-contig launch. Here the layout stays on the device too,
// and one compaction covers every contig:
//   1. count     per 4 KiB tile: contig starts and newlines;
//   2. (scans)   exclusive prefixes of both;
//   3. starts    the contig starts, in order (tiles without one return at once);
//   4. contigs   one wave per contig: its definition line end, sequence start
//                and base count, from the newline prefix (nothing is walked but
//                the definition line and two partial tiles);
//   5. (scan)    sequence offsets;
//   6. compact   per tile: a kept byte (a non-newline byte of a sequence) goes
//                to its rank among kept bytes, which is its contig's sequence
//                offset plus the non-newline bytes of that sequence before it.
//                A tile's kept bytes land contiguously, so they are staged in
//                LDS and written out as aligned dwords.
// The result equals the reference's wherever the reference's own compaction is
// in range (its uniform-wrap formula i - i / line_width and its byte-wise
// fallback both remove exactly the newlines). Traffic: the file is read twice
// (count, compact) and the sequence written once.
#include <hip/hip_runtime.h>

#include "text_internal.h"

namespace frhip {
namespace {

constexpr uint32_t THREADS = 256;
constexpr uint32_t PER = TX_TILE / THREADS;

__device__ inline uint32_t lane_id() { return threadIdx.x & 63; }

__device__ inline void load16(const uint8_t *buf, uint64_t len, uint64_t base, uint8_t b[16])
{
    if (base + 16 <= len && ((uintptr_t)(buf + base) & 15) == 0) {
        const uint4 v = *(const uint4 *)(buf + base);
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int k = 0; k < 16; ++k) b[k] = (uint8_t)(w[k / 4] >> (8 * (k % 4)));
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k) b[k] = base + k < len ? buf[base + k] : 0;
    }
}

__device__ inline uint32_t wave_incl_sum(uint32_t v)
{
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t x = __shfl_up(v, o, 64);
        if ((int)lane_id() >= o) v += x;
    }
    return v;
}

// Exclusive prefix over the 256 threads; *total: the sum.
__device__ inline uint32_t block_excl_sum(uint32_t v, uint32_t *lds, uint32_t *total)
{
    const uint32_t inc = wave_incl_sum(v);
    const uint32_t w = threadIdx.x / 64;
    if (lane_id() == 63) lds[w] = inc;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t k = 0; k < THREADS / 64; ++k) {
        if (k < w) before += lds[k];
        all += lds[k];
    }
    __syncthreads();
    if (total) *total = all;
    return before + inc - v;
}

__global__ __launch_bounds__(THREADS) void fasta_count_kernel(FastaArgs a)
{
    __shared__ uint32_t lds[THREADS / 64];
    const uint64_t t = blockIdx.x;
    const uint64_t base = t * TX_TILE + (uint64_t)threadIdx.x * PER;
    uint8_t b[16];
    load16(a.buf, a.len, base, b);
    const uint8_t prev = base > 0 && base - 1 < a.len ? a.buf[base - 1] : '\n';
    uint32_t gt = 0, nl = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (base + j >= a.len) continue;
        const uint8_t pc = j ? b[j - 1] : prev;
        gt += (b[j] == '>' && pc == '\n') ? 1u : 0u;
        nl += b[j] == '\n' ? 1u : 0u;
    }
    uint32_t tg, tn;
    block_excl_sum(gt, lds, &tg);
    block_excl_sum(nl, lds, &tn);
    if (threadIdx.x == 0) {
        a.t_gt[t] = tg;
        a.t_nl[t] = tn;
    }
}

__global__ __launch_bounds__(THREADS) void fasta_starts_kernel(FastaArgs a)
{
    __shared__ uint32_t lds[THREADS / 64];
    const uint64_t t = blockIdx.x;
    if (a.t_gt[t] == 0) return; // the whole workgroup leaves together
    const uint64_t base = t * TX_TILE + (uint64_t)threadIdx.x * PER;
    uint8_t b[16];
    load16(a.buf, a.len, base, b);
    const uint8_t prev = base > 0 && base - 1 < a.len ? a.buf[base - 1] : '\n';
    uint32_t gt = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint8_t pc = j ? b[j - 1] : prev;
        gt += (base + j < a.len && b[j] == '>' && pc == '\n') ? 1u : 0u;
    }
    uint64_t k = a.gt_base[t] + block_excl_sum(gt, lds, nullptr);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint8_t pc = j ? b[j - 1] : prev;
        if (base + j < a.len && b[j] == '>' && pc == '\n') a.starts[k++] = base + j;
    }
}

// Newlines before position x (x <= len), counted by one wave: the prefix up
// to x's tile plus the newlines of [tile start, x).
__device__ inline uint64_t wave_nl_before(const FastaArgs &a, uint64_t x)
{
    const uint64_t t = x / TX_TILE;
    const uint64_t t0 = t * TX_TILE;
    uint32_t n = 0;
    for (uint64_t p = t0 + lane_id(); p < x; p += 64) n += a.buf[p] == '\n' ? 1u : 0u;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    return a.nl_base[t] + n;
}

// One wave per contig start.
__global__ __launch_bounds__(256) void fasta_contig_kernel(FastaArgs a)
{
    const uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64;
    if (c >= a.n_gt) return; // whole waves
    const uint64_t p = a.starts[c];
    const uint64_t next = c + 1 < a.n_gt ? a.starts[c + 1] : a.len;
    // the definition line's newline: the first '\n' in (p, next)
    uint64_t def_end = ~0ull;
    for (uint64_t q = p + 1; q < next; q += 64) {
        const uint64_t r = q + lane_id();
        const uint64_t hit = __ballot(r < next && a.buf[r] == '\n');
        if (hit) {
            def_end = q + (uint64_t)__ffsll((long long)hit) - 1;
            break;
        }
    }
    if (def_end == ~0ull) {
        // only the last start can lack one: it is not a record
        if (lane_id() == 0) {
            a.seq_start[c] = next;
            a.nonnl0[c] = 0;
            a.seq_len[c] = 0;
            a.result[FA_T_VALID] = c;
        }
        return;
    }
    const uint64_t s0 = def_end + 1;
    const uint64_t nl0 = wave_nl_before(a, s0);
    const uint64_t nl1 = wave_nl_before(a, next);
    if (lane_id() == 0) {
        a.seq_start[c] = s0;
        a.nonnl0[c] = s0 - nl0;
        a.seq_len[c] = (next - s0) - (nl1 - nl0);
        if (c + 1 == a.n_gt) a.result[FA_T_VALID] = a.n_gt;
    }
}

// Kept bytes before x: the sequence offset of x's contig plus that sequence's
// bases before x (its definition line keeps none).
__device__ inline uint64_t kept_before(const FastaArgs &a, uint64_t n_valid, uint64_t total, uint64_t x,
                                       uint64_t nl_x)
{
    // last contig start <= x, over every start (an invalid last one included)
    if (a.n_gt == 0 || a.starts[0] > x) return 0;
    uint64_t lo = 0, hi = a.n_gt;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) / 2;
        if (a.starts[mid] <= x) lo = mid;
        else hi = mid;
    }
    if (lo >= n_valid) return total;
    if (x < a.seq_start[lo]) return a.seq_off[lo];
    return a.seq_off[lo] + (x - nl_x) - a.nonnl0[lo];
}

__global__ __launch_bounds__(THREADS) void fasta_compact_kernel(FastaArgs a)
{
    __shared__ uint32_t lds[THREADS / 64];
    __shared__ uint8_t stage[TX_TILE];
    __shared__ uint64_t tile_d0;
    const uint64_t t = blockIdx.x;
    const uint64_t n_valid = a.result[FA_T_VALID];
    const uint64_t total = a.seq_off[n_valid];
    const uint64_t tb = t * TX_TILE;
    const uint64_t base = tb + (uint64_t)threadIdx.x * PER;
    uint8_t b[16];
    load16(a.buf, a.len, base, b);
    uint32_t nl = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) nl += (base + j < a.len && b[j] == '\n') ? 1u : 0u;
    // newlines before this thread's first byte
    const uint64_t nl_base = a.nl_base[t] + block_excl_sum(nl, lds, nullptr);
    if (threadIdx.x == 0) tile_d0 = kept_before(a, n_valid, total, tb, a.nl_base[t]);
    // the contig of this thread's first byte, and where the next one starts
    uint64_t c = ~0ull, c_next = a.n_gt ? a.starts[0] : ~0ull;
    if (a.n_gt && a.starts[0] <= base) {
        uint64_t lo = 0, hi = a.n_gt;
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) / 2;
            if (a.starts[mid] <= base) lo = mid;
            else hi = mid;
        }
        c = lo;
        c_next = lo + 1 < a.n_gt ? a.starts[lo + 1] : ~0ull;
    }
    __syncthreads();
    const uint64_t d0 = tile_d0;
    uint64_t nl_x = nl_base;
    uint32_t kept = 0;
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const uint64_t p = base + j;
        if (p >= a.len) break;
        while (p >= c_next) {
            c = c == ~0ull ? 0 : c + 1;
            c_next = c + 1 < a.n_gt ? a.starts[c + 1] : ~0ull;
        }
        const uint8_t ch = b[j];
        if (ch == '\n') {
            ++nl_x;
            continue;
        }
        if (c == ~0ull || c >= n_valid || p < a.seq_start[c]) continue;
        const uint64_t dst = a.seq_off[c] + (p - nl_x) - a.nonnl0[c];
        stage[dst - d0] = ch;
        ++kept;
    }
    uint32_t k_all;
    block_excl_sum(kept, lds, &k_all);
    // write stage[0, k_all) to out[d0, d0 + k_all): bytes up to the first
    // dword boundary and after the last one, aligned dwords between
    const uint64_t e = d0 + k_all;
    uint64_t a0 = (d0 + 3) & ~3ull;
    uint64_t a1 = e & ~3ull;
    if (a0 > e) a0 = e;
    if (a1 < a0) a1 = a0;
    for (uint64_t q = d0 + threadIdx.x; q < a0; q += THREADS) a.out[q] = stage[q - d0];
    for (uint64_t q = a1 + threadIdx.x; q < e; q += THREADS) a.out[q] = stage[q - d0];
    for (uint64_t w = a0 + 4ull * threadIdx.x; w < a1; w += 4ull * THREADS) {
        const uint64_t o = w - d0;
        const uint32_t v = (uint32_t)stage[o] | ((uint32_t)stage[o + 1] << 8) | ((uint32_t)stage[o + 2] << 16) |
                           ((uint32_t)stage[o + 3] << 24);
        *(uint32_t *)(a.out + w) = v;
    }
}

} // namespace

hipError_t launch_fasta_count(const FastaArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(fasta_count_kernel, dim3((uint32_t)a.n_tiles), dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fasta_starts(const FastaArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(fasta_starts_kernel, dim3((uint32_t)a.n_tiles), dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fasta_contigs(const FastaArgs &a, hipStream_t s)
{
    const uint64_t threads = a.n_gt * 64;
    hipLaunchKernelGGL(fasta_contig_kernel, dim3((uint32_t)((threads + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fasta_compact(const FastaArgs &a, hipStream_t s)
{
    hipLaunchKernelGGL(fasta_compact_kernel, dim3((uint32_t)a.n_tiles), dim3(THREADS), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
