// region_internal.h -- device side of region queries over a served batch.
#ifndef FRHIP_REGION_INTERNAL_H
#define FRHIP_REGION_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pipeline_internal.h"

namespace frhip {

// A region as the filter reads it: reference id, and for a bounded region the
// 1-based closed interval [start, end] (an open bound resolved to 1 or to
// INT64_MAX).
struct RegionDev {
    int32_t ref_id;
    uint32_t bounded;
    uint64_t start, end;
};
static_assert(sizeof(RegionDev) == 24, "region table entry");

struct RegionFilterArgs {
    const uint8_t *buf;             // the batch's inflated bytes
    const uint64_t *record_offsets; // row -> record offset in buf
    const int32_t *ref_id, *pos;    // columns
    const uint32_t *cigar_start, *seq_start;
    const uint64_t *seg_base;       // ns + 1: rows of segment sq are [seg_base[sq], seg_base[sq + 1])
    uint32_t ns;
    const uint32_t *seg_region;     // region of each segment
    const RegionDev *regions;
    // scratch
    uint64_t *mask;                 // 4 ballot words per tile of 256 rows
    uint32_t *tile_cnt;             // kept rows per tile
    uint64_t *tile_base;            // exclusive prefix of tile_cnt, tiles + 1
    uint64_t *n_tiles;              // tiles the batch's rows fill
    unsigned long long *seg_bad;    // per segment: first row failing the check (init ~0)
    uint64_t *sel_first;            // ns + 1: selection index of each segment's first row
    // output: kept rows in row order
    uint64_t *sel;
};

struct RegionScratchLayout {
    uint64_t o_mask, o_cnt, o_base, o_ntiles, o_bad, o_first, bytes;
};

RegionScratchLayout region_scratch_layout(uint64_t cap_rows, uint64_t ns);
void region_bind_scratch(RegionFilterArgs &a, uint8_t *s, const RegionScratchLayout &L);
// mark, scan, compact, segment starts; seg_bad must hold ~0 beforehand
hipError_t launch_region_filter(const RegionFilterArgs &a, uint64_t cap_rows, hipStream_t s);

} // namespace frhip
#endif
