// region.hip -- keep the records of a region query that overlap the region.
//
// After a served batch has decoded the records of every chunk a region reads,
// each record is tested on the device against its region, as
// crates/fritillaria-bam/src/io/reader/query.rs:104-132 does per record on the
// host: same reference, and for a bounded region an alignment [pos + 1,
// pos + span] (span from the raw CIGAR's M/D/N/=/X lengths, at least 1) that
// overlaps the region's 1-based closed interval
// (crates/fritillaria-sam/src/alignment/record.rs:108-135,
// crates/fritillaria-core/src/region/interval.rs:96-108). A record whose
// reference id is below -1, or, on the region's reference of a bounded
// region, whose position is below -1 or whose CIGAR holds an op kind above 8,
// is where the reference's check fails: the first such row of each segment is
// recorded. The kept rows are compacted in row order into one selection, so a
// region's rows are one contiguous stretch of it.
//
// Rows are taken in tiles of 256 (a workgroup, four waves): one pass marks
// (a ballot word per wave, a count per tile), a one-workgroup scan turns the
// counts into tile bases, one pass writes the rows, and one small pass gives
// each segment the selection index of its first row. Traffic per record: the
// two columns and the record offset (16 bytes), its CIGAR when the region is
// bounded (4 bytes per op), and 8 bytes per kept row.
#include <hip/hip_runtime.h>

#include "region_internal.h"

namespace frhip {
namespace {

constexpr uint32_t TILE = 256;

__device__ inline uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

// Largest segment sq in [0, ns) with seg_base[sq] <= row (row < seg_base[ns]).
__device__ inline uint32_t segment_of(const uint64_t *seg_base, uint32_t ns, uint64_t row)
{
    uint32_t lo = 0, hi = ns; // answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (seg_base[mid] <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline uint64_t lanes_below(uint64_t ballot)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(ballot >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)ballot, 0u));
}

__global__ __launch_bounds__(TILE) void region_mark_kernel(RegionFilterArgs a)
{
    __shared__ uint32_t wave_cnt[TILE / 64];
    const uint64_t n_rows = a.seg_base[a.ns];
    const uint64_t tiles = (n_rows + TILE - 1) / TILE;
    if (blockIdx.x == 0 && threadIdx.x == 0) *a.n_tiles = tiles;
    const uint32_t w = threadIdx.x / 64;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t row = t * TILE + threadIdx.x;
        bool keep = false;
        if (row < n_rows) {
            const uint32_t sq = segment_of(a.seg_base, a.ns, row);
            const RegionDev R = a.regions[a.seg_region[sq]];
            const int32_t rid = a.ref_id[row];
            const int32_t p = a.pos[row];
            bool bad = rid < -1;
            if (!bad && rid == R.ref_id) {
                if (!R.bounded) {
                    keep = true;
                } else if (p < -1) {
                    bad = true;
                } else if (p >= 0) {
                    const uint8_t *cig = a.buf + a.record_offsets[row] + a.cigar_start[row];
                    const uint32_t n_ops = (a.seq_start[row] - a.cigar_start[row]) / 4;
                    uint64_t span = 0;
                    for (uint32_t k = 0; k < n_ops; ++k) {
                        const uint32_t op = rd32(cig + 4 * k);
                        const uint32_t kind = op & 0xF;
                        if (kind > 8) {
                            bad = true;
                            break;
                        }
                        // M, D, N, =, X consume the reference
                        if ((0x18Du >> kind) & 1u) span += op >> 4;
                    }
                    const uint64_t s0 = (uint64_t)p + 1;
                    const uint64_t e0 = span ? s0 + span - 1 : s0;
                    keep = !bad && s0 <= R.end && R.start <= e0;
                }
            }
            if (bad) {
                atomicMin(&a.seg_bad[sq], (unsigned long long)row);
                keep = false;
            }
        }
        const uint64_t b = __ballot(keep);
        if ((threadIdx.x & 63) == 0) {
            a.mask[t * 4 + w] = b;
            wave_cnt[w] = (uint32_t)__popcll(b);
        }
        __syncthreads();
        if (threadIdx.x == 0) a.tile_cnt[t] = wave_cnt[0] + wave_cnt[1] + wave_cnt[2] + wave_cnt[3];
        __syncthreads();
    }
}

__global__ __launch_bounds__(TILE) void region_compact_kernel(RegionFilterArgs a)
{
    const uint64_t tiles = *a.n_tiles;
    const uint32_t w = threadIdx.x / 64;
    for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
        const uint64_t b = a.mask[t * 4 + w];
        if (!((b >> (threadIdx.x & 63)) & 1ull)) continue;
        uint64_t before = a.tile_base[t];
        for (uint32_t j = 0; j < w; ++j) before += (uint64_t)__popcll(a.mask[t * 4 + j]);
        a.sel[before + lanes_below(b)] = t * TILE + threadIdx.x;
    }
}

// sel_first[sq] for sq in [0, ns]: the selection index of segment sq's first
// row (for an empty segment and for ns, the count kept before the row there).
__global__ void region_seg_first_kernel(RegionFilterArgs a)
{
    const uint32_t sq = blockIdx.x * blockDim.x + threadIdx.x;
    if (sq > a.ns) return;
    const uint64_t n_rows = a.seg_base[a.ns];
    const uint64_t tiles = *a.n_tiles;
    const uint64_t r = a.seg_base[sq];
    uint64_t v;
    if (r >= n_rows) {
        v = a.tile_base[tiles];
    } else {
        const uint64_t t = r / TILE;
        const uint32_t in = (uint32_t)(r % TILE), w = in / 64, bit = in % 64;
        v = a.tile_base[t];
        for (uint32_t j = 0; j < w; ++j) v += (uint64_t)__popcll(a.mask[t * 4 + j]);
        if (bit) v += (uint64_t)__popcll(a.mask[t * 4 + w] & ((1ull << bit) - 1));
    }
    a.sel_first[sq] = v;
}

} // namespace

RegionScratchLayout region_scratch_layout(uint64_t cap_rows, uint64_t ns)
{
    RegionScratchLayout L{};
    const uint64_t tiles = (cap_rows + TILE - 1) / TILE;
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = (p + (b ? b : 1) + 255) & ~255ull; return o; };
    L.o_mask = take(tiles * 32);
    L.o_cnt = take(tiles * 4);
    L.o_base = take((tiles + 1) * 8);
    L.o_ntiles = take(8);
    L.o_bad = take(ns * 8);
    L.o_first = take((ns + 1) * 8);
    L.bytes = p;
    return L;
}

void region_bind_scratch(RegionFilterArgs &a, uint8_t *s, const RegionScratchLayout &L)
{
    a.mask = (uint64_t *)(s + L.o_mask);
    a.tile_cnt = (uint32_t *)(s + L.o_cnt);
    a.tile_base = (uint64_t *)(s + L.o_base);
    a.n_tiles = (uint64_t *)(s + L.o_ntiles);
    a.seg_bad = (unsigned long long *)(s + L.o_bad);
    a.sel_first = (uint64_t *)(s + L.o_first);
}

hipError_t launch_region_filter(const RegionFilterArgs &a, uint64_t cap_rows, hipStream_t s)
{
    uint64_t tiles = (cap_rows + TILE - 1) / TILE;
    uint32_t grid = (uint32_t)(tiles < 8192 ? (tiles ? tiles : 1) : 8192);
    hipLaunchKernelGGL(region_mark_kernel, dim3(grid), dim3(TILE), 0, s, a);
    hipError_t r = hipGetLastError();
    if (r != hipSuccess) return r;
    r = launch_scan_u32_u64(a.tile_cnt, a.tile_base, a.n_tiles, 0, s);
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(region_compact_kernel, dim3(grid), dim3(TILE), 0, s, a);
    r = hipGetLastError();
    if (r != hipSuccess) return r;
    hipLaunchKernelGGL(region_seg_first_kernel, dim3((a.ns + 1 + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

} // namespace frhip
