// pipeline.hip -- kernels of the batched, device-resident path: BGZF member
// discovery on the device, batch verification, the BAM header walk, a
// single-workgroup exclusive scan and a fill kernel (graph-safe clearing).
//
// Member discovery restates crates/fritillaria-bgzf/src/discover.rs:51-113 and
// block.rs:99-164 as a sieve: every byte offset whose bytes parse as a BGZF
// header (magic, CM=8, FEXTRA, a BC subfield, a block size that fits its own
// framing) is a candidate, compacted in file order. The candidates are the
// true member chain exactly when the first is at 0, each one's block size lands
// on the next, and the last lands on the end of the file; the host checks that
// proof once and otherwise falls back to its serial walk.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "frhip_internal.h"
#include "pipeline_internal.h"

namespace frhip {
namespace {

__device__ __forceinline__ uint32_t rd16(const uint8_t *p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8); }
__device__ __forceinline__ uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// block.rs:99-164 on the bytes at p; returns block size (0 if not a header).
// The usual layout (one extra subfield, BC, so an 18-byte header) is checked
// from 18 bytes loaded together; any other goes through the general walk.
__device__ __forceinline__ uint32_t header_at(const uint8_t *f, uint64_t len, uint64_t p, uint32_t *hsize)
{
    if (p + 12 > len) return 0;
    const uint8_t *b = f + p;
    if (p + 18 <= len) {
        uint32_t h[18];
#pragma unroll
        for (int i = 0; i < 18; ++i) h[i] = b[i];
        // all 18 loads issue before the first check (one wait, not one per byte)
        asm volatile("" ::"v"(h[0]), "v"(h[1]), "v"(h[2]), "v"(h[3]), "v"(h[4]), "v"(h[5]), "v"(h[6]), "v"(h[7]),
                     "v"(h[8]), "v"(h[9]), "v"(h[10]), "v"(h[11]), "v"(h[12]), "v"(h[13]), "v"(h[14]), "v"(h[15]),
                     "v"(h[16]), "v"(h[17]));
        if (h[0] != 0x1f || h[1] != 0x8b || h[2] != 8 || (h[3] & 4) == 0) return 0;
        if ((h[10] | h[11] << 8) == 6 && h[12] == 'B' && h[13] == 'C' && (h[14] | h[15] << 8) == 2) {
            const uint32_t bs = (h[16] | h[17] << 8) + 1;
            if (bs < 18 + 8 || p + bs > len) return 0;
            *hsize = 18;
            return bs;
        }
    }
    if (b[0] != 0x1f || b[1] != 0x8b || b[2] != 8 || (b[3] & 4) == 0) return 0;
    uint32_t xlen = rd16(b + 10);
    uint64_t hs = 12 + (uint64_t)xlen;
    if (p + hs > len) return 0;
    uint32_t off = 0;
    for (int k = 0; k < 64 && off + 4 <= xlen; ++k) {
        uint32_t slen = rd16(b + 12 + off + 2);
        if (off + 4 + slen > xlen) return 0;
        if (b[12 + off] == 'B' && b[12 + off + 1] == 'C') {
            if (slen != 2) return 0;
            uint32_t bs = rd16(b + 12 + off + 4) + 1;
            if (bs < hs + 8) return 0;
            if (p + bs > len) return 0;
            *hsize = (uint32_t)hs;
            return bs;
        }
        off += 4 + slen;
    }
    return 0;
}

constexpr uint32_t SIEVE_TILE = 16384;
constexpr uint32_t SIEVE_THREADS = 256;
constexpr uint32_t SIEVE_CHUNK = 16; // bytes per thread per pass over a 4 KiB stretch

// Bit 7 of each byte of w set where that byte equals the byte of x4 (x4 holds
// it four times; exact, no carries between bytes).
__device__ __forceinline__ uint32_t bytes_eq(uint32_t w, uint32_t x4)
{
    const uint32_t t = w ^ x4;
    return ~(((t & 0x7f7f7f7fu) + 0x7f7f7f7fu) | t | 0x7f7f7f7fu);
}
// Offsets k of a 16-byte chunk (words v) where the gzip magic 1f 8b starts, as
// a 16-bit mask: each word against 0x1f, the same word shifted one byte (the
// bytes after, via a funnel shift) against 0x8b. The chunk's last byte keeps a
// 0x1f alone (its successor is in the next chunk): a superset, header_at
// decides.
__device__ __forceinline__ uint32_t magic_nibbles(const uint4 v)
{
    auto nib = [](uint32_t w, uint32_t after) {
        const uint32_t m = bytes_eq(w, 0x1f1f1f1fu) & bytes_eq(after, 0x8b8b8b8bu);
        return ((m >> 7) & 1u) | ((m >> 14) & 2u) | ((m >> 21) & 4u) | ((m >> 28) & 8u);
    };
    return nib(v.x, __builtin_amdgcn_alignbit(v.y, v.x, 8)) | (nib(v.y, __builtin_amdgcn_alignbit(v.z, v.y, 8)) << 4) |
           (nib(v.z, __builtin_amdgcn_alignbit(v.w, v.z, 8)) << 8) |
           (nib(v.w, __builtin_amdgcn_alignbit(0x8b8b8b8bu, v.w, 8)) << 12);
}

// Offsets p .. p+15 (p a multiple of 16) where a header may start, as a 16-bit
// mask: one aligned 16-byte load through magic_nibbles (bytes past the end of
// the file are read one at a time and checked for 0x1f). Only these offsets go
// through header_at.
__device__ __forceinline__ uint32_t magic_mask(const uint8_t *f, uint64_t len, uint64_t p)
{
    if (p >= len) return 0;
    if (p + 16 <= len) {
        return magic_nibbles(*reinterpret_cast<const uint4 *>(f + p));
    }
    uint32_t m = 0;
    for (uint32_t j = 0; p + j < len && j < 16; ++j) m |= (uint32_t)(f[p + j] == 0x1f) << j;
    return m;
}

// magic_mask for a thread's NK chunks of the tile at base (chunk i at
// base + 16 * thread + 4096 * i): for a tile wholly inside the file, the NK
// loads are all in flight before the first is used.
template <uint32_t NK>
__device__ __forceinline__ void tile_masks(const uint8_t *f, uint64_t len, uint64_t base, uint32_t *mk)
{
    const uint64_t p0 = base + threadIdx.x * SIEVE_CHUNK;
    if (base + SIEVE_TILE <= len) {
        uint4 v[NK];
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) v[i] = *reinterpret_cast<const uint4 *>(f + p0 + i * SIEVE_THREADS * SIEVE_CHUNK);
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) mk[i] = magic_nibbles(v[i]);
    } else {
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) mk[i] = magic_mask(f, len, p0 + i * SIEVE_THREADS * SIEVE_CHUNK);
    }
}

// Pass 1: candidates per tile.
__global__ __launch_bounds__(SIEVE_THREADS) void sieve_count_kernel(const uint8_t *f, uint64_t len, uint32_t *counts,
                                                                    uint64_t n_tiles)
{
    __shared__ uint32_t wsum[SIEVE_THREADS / 64];
    for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint64_t base = t * SIEVE_TILE;
        uint32_t c = 0;
        constexpr uint32_t NK = SIEVE_TILE / (SIEVE_THREADS * SIEVE_CHUNK);
        uint32_t mk[NK];
        tile_masks<NK>(f, len, base, mk);
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) {
            const uint64_t p = base + threadIdx.x * SIEVE_CHUNK + i * SIEVE_THREADS * SIEVE_CHUNK;
            uint32_t m = mk[i];
            while (m) {
                const uint32_t j = (uint32_t)__ffs(m) - 1;
                m &= m - 1;
                uint32_t hs;
                if (header_at(f, len, p + j, &hs)) ++c;
            }
        }
        for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
        if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t s = 0;
            for (uint32_t w = 0; w < SIEVE_THREADS / 64; ++w) s += wsum[w];
            counts[t] = s;
        }
        __syncthreads();
    }
}

// Pass 2: candidates written in file order at the tile's scanned base. Each
// thread takes 16 consecutive offsets; a block-wide exclusive scan of the
// per-thread counts places them.
__global__ __launch_bounds__(SIEVE_THREADS) void sieve_write_kernel(const uint8_t *f, uint64_t len,
                                                                    const uint64_t *tile_base, uint64_t n_tiles,
                                                                    uint64_t *cand, uint32_t *cand_bs,
                                                                    uint32_t *cand_hs, uint64_t cap)
{
    constexpr uint32_t NK = SIEVE_TILE / (SIEVE_THREADS * SIEVE_CHUNK);
    constexpr uint32_t NW = SIEVE_THREADS / 64;
    __shared__ uint32_t wcnt[NK][NW];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
        const uint64_t base = t * SIEVE_TILE;
        const uint64_t out = tile_base[t];
        uint32_t mk[NK];
        tile_masks<NK>(f, len, base, mk);
        // the headers among each chunk's candidates, and their counts scanned
        // over the wave (file order: chunk, then thread)
        uint32_t vm[NK], c[NK], inc[NK];
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) {
            const uint64_t p = base + i * SIEVE_THREADS * SIEVE_CHUNK + threadIdx.x * SIEVE_CHUNK;
            uint32_t m = mk[i];
            vm[i] = 0;
            while (m) {
                const uint32_t j = (uint32_t)__ffs(m) - 1;
                m &= m - 1;
                uint32_t hs;
                if (header_at(f, len, p + j, &hs)) vm[i] |= 1u << j;
            }
            c[i] = (uint32_t)__popc(vm[i]);
            inc[i] = c[i];
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(inc[i], d, 64);
                if ((int)lane >= d) inc[i] += y;
            }
            if (lane == 63) wcnt[i][w] = inc[i];
        }
        __syncthreads();
        uint64_t before = out;
#pragma unroll
        for (uint32_t i = 0; i < NK; ++i) {
            uint32_t wb = 0, wt = 0;
#pragma unroll
            for (uint32_t v = 0; v < NW; ++v) {
                wb += v < w ? wcnt[i][v] : 0u;
                wt += wcnt[i][v];
            }
            const uint64_t p = base + i * SIEVE_THREADS * SIEVE_CHUNK + threadIdx.x * SIEVE_CHUNK;
            uint64_t at = before + wb + inc[i] - c[i];
            uint32_t m = vm[i];
            while (m) {
                const uint32_t j = (uint32_t)__ffs(m) - 1;
                m &= m - 1;
                uint32_t hs = 0;
                const uint32_t bs = header_at(f, len, p + j, &hs);
                if (at < cap) {
                    cand[at] = p + j;
                    cand_bs[at] = bs;
                    cand_hs[at] = hs;
                }
                ++at;
            }
            before += wt;
        }
        __syncthreads();
    }
}

// The tiling proof and the spans: member i is valid when it starts where
// member i-1 ends (0 for the first) and the last ends at the file end.
__global__ void spans_kernel(const uint8_t *f, uint64_t len, const uint64_t *cand, const uint32_t *cand_bs,
                             const uint32_t *cand_hs, const uint64_t *n_cand, uint64_t cap, PipeMembers m,
                             uint64_t *bad)
{
    uint64_t n = *n_cand;
    if (n > cap) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd((unsigned long long *)bad, 1ull);
        return;
    }
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t p = cand[i];
        uint64_t end = p + cand_bs[i];
        bool ok = (i == 0 ? p == 0 : true) && (i + 1 < n ? end == cand[i + 1] : end == len);
        if (!ok) atomicAdd((unsigned long long *)bad, 1ull);
        const uint8_t *b = f + p;
        uint32_t bs = cand_bs[i], hs = cand_hs[i];
        uint32_t isz = rd32(b + bs - 4);
        if (isz > 65536) atomicAdd((unsigned long long *)bad, 1ull);
        m.coff[i] = p;
        m.in_off[i] = p + hs;
        m.in_len[i] = bs - hs - 8;
        m.isize[i] = isz;
        m.crc_expect[i] = rd32(b + bs - 8);
    }
}

// Exclusive scan in three passes, no scratch beyond the output: blocks of
// SCAN_B elements (256 threads x 8) first park their sum in their own first
// output slot; one workgroup scans those slots in place (and writes the total
// to out[n]); then each block writes its elements' exclusive prefixes from
// that base. Grids are fixed and stride over blocks, so a count that is only
// known on the device (n_dev) needs no host round trip. in and out must not
// alias.
constexpr uint32_t SCAN_T = 256, SCAN_PER = 8, SCAN_B = SCAN_T * SCAN_PER;

template <typename T> __device__ inline T scan_wave_incl(T v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        T x = __shfl_up(v, o, 64);
        if (lane >= o) v += x;
    }
    return v;
}

// Exclusive prefix of v over the workgroup's NW waves; *total: the sum.
template <uint32_t NW, typename T> __device__ inline T scan_block_excl(T v, T *lds, T *total)
{
    const T inc = scan_wave_incl(v);
    const uint32_t w = threadIdx.x / 64;
    if ((threadIdx.x & 63) == 63) lds[w] = inc;
    __syncthreads();
    T before = 0, all = 0;
    for (uint32_t k = 0; k < NW; ++k) {
        if (k < w) before += lds[k];
        all += lds[k];
    }
    __syncthreads();
    *total = all;
    return before + inc - v;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_T) void scan_reduce_kernel(const TI *in, TO *out, const uint64_t *n_dev,
                                                             uint64_t n_host)
{
    __shared__ TO lds[SCAN_T / 64];
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t nb = (n + SCAN_B - 1) / SCAN_B;
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t i0 = b * SCAN_B + (uint64_t)threadIdx.x * SCAN_PER;
        TO sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER; ++k)
            if (i0 + k < n) sum += (TO)in[i0 + k];
        TO all;
        scan_block_excl<SCAN_T / 64>(sum, lds, &all);
        if (threadIdx.x == 0) out[b * SCAN_B] = all;
    }
}

template <typename TO>
__global__ __launch_bounds__(1024) void scan_blocks_kernel(TO *out, const uint64_t *n_dev, uint64_t n_host)
{
    __shared__ TO lds[16];
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t nb = (n + SCAN_B - 1) / SCAN_B;
    TO run = 0;
    for (uint64_t c0 = 0; c0 < nb; c0 += 1024) {
        const uint64_t b = c0 + threadIdx.x;
        const TO v = b < nb ? out[b * SCAN_B] : (TO)0;
        TO all;
        const TO ex = scan_block_excl<16>(v, lds, &all);
        if (b < nb) out[b * SCAN_B] = run + ex;
        run += all;
    }
    if (threadIdx.x == 0) out[n] = run;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(SCAN_T) void scan_down_kernel(const TI *in, TO *out, const uint64_t *n_dev,
                                                           uint64_t n_host)
{
    __shared__ TO lds[SCAN_T / 64];
    __shared__ TO base_s;
    const uint64_t n = n_dev ? *n_dev : n_host;
    const uint64_t nb = (n + SCAN_B - 1) / SCAN_B;
    for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        if (threadIdx.x == 0) base_s = out[b * SCAN_B];
        const uint64_t i0 = b * SCAN_B + (uint64_t)threadIdx.x * SCAN_PER;
        TO v[SCAN_PER];
        TO sum = 0;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER; ++k) {
            v[k] = i0 + k < n ? (TO)in[i0 + k] : (TO)0;
            sum += v[k];
        }
        TO all;
        TO run = scan_block_excl<SCAN_T / 64>(sum, lds, &all); // its barriers also publish base_s
        run += base_s;
#pragma unroll
        for (uint32_t k = 0; k < SCAN_PER; ++k) {
            if (i0 + k < n) out[i0 + k] = run;
            run += v[k];
        }
        __syncthreads(); // base_s is rewritten for the next block
    }
}

template <typename TI, typename TO>
hipError_t launch_scan(const TI *in, TO *out, const uint64_t *n_dev, uint64_t n_host, hipStream_t s)
{
    uint64_t grid = 1024;
    if (!n_dev) {
        grid = (n_host + SCAN_B - 1) / SCAN_B;
        if (grid > 1024) grid = 1024;
        if (grid == 0) grid = 1;
    }
    hipLaunchKernelGGL((scan_reduce_kernel<TI, TO>), dim3((uint32_t)grid), dim3(SCAN_T), 0, s, in, out, n_dev, n_host);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL((scan_blocks_kernel<TO>), dim3(1), dim3(1024), 0, s, out, n_dev, n_host);
    r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL((scan_down_kernel<TI, TO>), dim3((uint32_t)grid), dim3(SCAN_T), 0, s, in, out, n_dev, n_host);
    return hipGetLastError();
}

// First failing member in file order, with the check that failed
// (status, then size, then CRC: crates/fritillaria-cuda/src/backend.rs:221-256).
__global__ void verify_kernel(const uint32_t *status, const uint32_t *produced, const uint32_t *crc, PipeMembers m,
                              const uint64_t *n_dev, unsigned long long *first_bad)
{
    uint64_t n = *n_dev;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        bool bad = status[i] != 0 || produced[i] != m.isize[i] || crc[i] != m.crc_expect[i];
        if (bad) atomicMin(first_bad, (unsigned long long)i);
    }
}

// crates/fritillaria-bam/src/columnar/header.rs:43-103 on the inflated buffer.
__global__ void bam_header_kernel(const uint8_t *u, const uint64_t *u_len, uint64_t *records_start, uint64_t *status)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t len = *u_len;
    uint64_t st = 0, pos = 0;
    if (len < 12 || u[0] != 'B' || u[1] != 'A' || u[2] != 'M' || u[3] != 1) {
        st = 1;
    } else {
        int32_t l_text = (int32_t)rd32(u + 4);
        uint64_t te = 8 + (uint64_t)(l_text < 0 ? 0 : l_text);
        if (l_text < 0 || te + 4 > len) {
            st = 1;
        } else {
            int32_t nref = (int32_t)rd32(u + te);
            pos = te + 4;
            if (nref < 0) st = 1;
            for (int32_t i = 0; i < nref && !st; ++i) {
                if (pos + 4 > len) { st = 1; break; }
                int32_t l_name = (int32_t)rd32(u + pos);
                if (l_name < 0 || pos + 4 + (uint64_t)l_name + 4 > len) { st = 1; break; }
                pos += 4 + (uint64_t)l_name + 4;
            }
        }
    }
    *records_start = st ? 0 : pos;
    *status = st;
}

// Plain device copy: 16-byte words, grid-stride, byte tail.
__global__ __launch_bounds__(256) void copy_kernel(uint8_t *dst, const uint8_t *src, uint64_t n)
{
    const uint64_t words = n >> 4;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride)
        reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
    for (uint64_t i = (words << 4) + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

__global__ void fill_u64_kernel(uint64_t *p, uint64_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// When the discovered members' inflated total is not the size the output was
// built for (a re-uploaded file whose member table differs), every member's
// output extent becomes empty: the inflate then writes nothing and the run
// fails verification instead of writing past the output.
__global__ void guard_total_kernel(uint64_t *out_off, uint32_t *isize, uint64_t n, uint64_t expect,
                                   uint64_t *bad)
{
    if (out_off[n] == expect) return;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
        out_off[i] = 0;
        isize[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        out_off[n] = 0;
        *bad = 1;
    }
}

} // namespace

uint64_t sieve_tiles(uint64_t len) { return (len + SIEVE_TILE - 1) / SIEVE_TILE; }

hipError_t launch_sieve_count(const uint8_t *f, uint64_t len, uint32_t *counts, hipStream_t s)
{
    uint64_t nt = sieve_tiles(len);
    if (nt == 0) return hipSuccess;
    uint32_t grid = (uint32_t)(nt < 65536 ? nt : 65536);
    hipLaunchKernelGGL(sieve_count_kernel, dim3(grid), dim3(SIEVE_THREADS), 0, s, f, len, counts, nt);
    return hipGetLastError();
}

hipError_t launch_sieve_write(const uint8_t *f, uint64_t len, const uint64_t *tile_base, uint64_t *cand,
                              uint32_t *cand_bs, uint32_t *cand_hs, uint64_t cap, hipStream_t s)
{
    uint64_t nt = sieve_tiles(len);
    if (nt == 0) return hipSuccess;
    uint32_t grid = (uint32_t)(nt < 65536 ? nt : 65536);
    hipLaunchKernelGGL(sieve_write_kernel, dim3(grid), dim3(SIEVE_THREADS), 0, s, f, len, tile_base, nt, cand, cand_bs,
                       cand_hs, cap);
    return hipGetLastError();
}

hipError_t launch_spans(const uint8_t *f, uint64_t len, const uint64_t *cand, const uint32_t *cand_bs,
                        const uint32_t *cand_hs, const uint64_t *n_cand, uint64_t cap, const PipeMembers &m,
                        uint64_t *bad, hipStream_t s)
{
    uint32_t grid = (uint32_t)((cap + 255) / 256);
    if (grid == 0) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(spans_kernel, dim3(grid), dim3(256), 0, s, f, len, cand, cand_bs, cand_hs, n_cand, cap, m, bad);
    return hipGetLastError();
}

hipError_t launch_scan_u32_u64(const uint32_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_host,
                               hipStream_t s)
{
    return launch_scan(in, out, n_dev, n_host, s);
}

hipError_t launch_scan_u64_u64(const uint64_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_host,
                               hipStream_t s)
{
    return launch_scan(in, out, n_dev, n_host, s);
}

hipError_t launch_verify(const uint32_t *status, const uint32_t *produced, const uint32_t *crc, const PipeMembers &m,
                         const uint64_t *n_dev, uint64_t cap, unsigned long long *first_bad, hipStream_t s)
{
    uint32_t grid = (uint32_t)((cap + 255) / 256);
    if (grid == 0) grid = 1;
    if (grid > 4096) grid = 4096;
    hipLaunchKernelGGL(verify_kernel, dim3(grid), dim3(256), 0, s, status, produced, crc, m, n_dev, first_bad);
    return hipGetLastError();
}

hipError_t launch_bam_header(const uint8_t *u, const uint64_t *u_len, uint64_t *records_start, uint64_t *status,
                             hipStream_t s)
{
    hipLaunchKernelGGL(bam_header_kernel, dim3(1), dim3(64), 0, s, u, u_len, records_start, status);
    return hipGetLastError();
}

hipError_t launch_copy(uint8_t *dst, const uint8_t *src, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint64_t grid = (n / 16 + 255) / 256;
    if (grid > 8192) grid = 8192;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)grid), dim3(256), 0, s, dst, src, n);
    return hipGetLastError();
}

hipError_t launch_guard_total(uint64_t *out_off, uint32_t *isize, uint64_t n, uint64_t expect, uint64_t *bad,
                              hipStream_t s)
{
    uint64_t grid = (n + 255) / 256;
    if (grid > 1024) grid = 1024;
    if (grid == 0) grid = 1;
    hipLaunchKernelGGL(guard_total_kernel, dim3((uint32_t)grid), dim3(256), 0, s, out_off, isize, n, expect, bad);
    return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    uint32_t grid = (uint32_t)((n + 255) / 256);
    if (grid > 1024) grid = 1024;
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid), dim3(256), 0, s, p, v, n);
    return hipGetLastError();
}

} // namespace frhip
