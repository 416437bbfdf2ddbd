// fastq_internal.h -- device arrays and launchers of the FASTQ boundary scan.
#ifndef FRHIP_FASTQ_INTERNAL_H
#define FRHIP_FASTQ_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

constexpr uint32_t FQ_TILE = 65536;  // bytes per workgroup in the newline passes
constexpr uint32_t FQ_SEG = 256;     // lines per segment of the parallel walk
constexpr uint64_t FQ_PAD = 64 * FQ_SEG; // line-type array granule (one segment workgroup)
constexpr uint32_t FQ_EXIT_DONE = 4; // the walk ended inside the segment (or before it)

enum : uint8_t { FQ_BLANK = 0, FQ_RECORD = 1, FQ_STOP = 2, FQ_END = 3 };
enum { FQ_T_RECORDS = 0, FQ_T_TAIL = 1, FQ_T_STATUS = 2, FQ_T_SURVIVORS = 3, FQ_T_N = 4 };

struct FastqArgs {
    const uint8_t *buf;
    uint64_t start, len;
    uint64_t first_tile, n_tiles; // 64 KiB tiles, aligned to absolute offsets
    uint32_t *tile_counts;        // n_tiles
    uint64_t *tile_base;          // n_tiles + 1 (exclusive scan, total last)
    uint64_t *nl;                 // n_nl newline positions, ascending
    uint64_t n_nl;
    uint64_t n_lines;             // n_nl, plus one when the batch does not end in '\n'
    uint64_t n_lines_pad;         // >= n_lines + 4, a multiple of FQ_PAD
    uint8_t *type;                // n_lines_pad line types
    uint64_t n_seg;               // n_lines_pad / FQ_SEG
    uint64_t *seg_f;              // per segment: 4 x 16-bit (exit | count << 3)
    uint8_t *seg_state;           // true entry of each segment (>= 4: the walk ended before it)
    uint64_t *seg_base;           // record index of each segment's first record
    uint64_t *totals;             // FQ_T_*
    uint64_t cap;                 // record capacity of the columns
    uint64_t *record_offsets;
    uint32_t *sequence_start, *plus_start, *quality_start, *sequence_len, *record_end;
};

uint64_t fastq_tiles(uint64_t start, uint64_t len);
// newline counts per tile, then their exclusive scan (total at tile_base[n_tiles])
hipError_t launch_fastq_index(const FastqArgs &a, hipStream_t s);
// newline positions, line types, segment functions, their composition
hipError_t launch_fastq_lines(const FastqArgs &a, hipStream_t s);
// record offsets, line bounds, tail and status
hipError_t launch_fastq_emit(const FastqArgs &a, hipStream_t s);

hipError_t launch_scan_u32_u64(const uint32_t *in, uint64_t *out, const uint64_t *n_dev, uint64_t n_host,
                               hipStream_t s);
hipError_t launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);

} // namespace frhip
#endif
