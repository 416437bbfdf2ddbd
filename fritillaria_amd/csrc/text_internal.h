// text_internal.h -- device arrays and launchers of text framing and FASTA compaction.
#ifndef FRHIP_TEXT_INTERNAL_H
#define FRHIP_TEXT_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

// Both paths take the batch in tiles of 4096 bytes (256 threads x 16 bytes),
// tile t covering [t * 4096, (t + 1) * 4096) of the buffer.
constexpr uint32_t TX_TILE = 4096;
inline uint64_t tx_tiles(uint64_t len) { return (len + TX_TILE - 1) / TX_TILE; }
inline uint64_t text_groups(uint64_t tiles) { return (tiles + 1023) / 1024; }

enum { TX_NONE = 0, TX_RECORD = 1, TX_HEADER = 2 };

// Totals of a text scan (one workgroup fills them from the tile counts).
enum { TX_T_RECORDS = 0, TX_T_HEADERS = 1, TX_T_TABS = 2, TX_T_LAST_LS = 3, TX_T_LAST_KIND = 4, TX_T_N = 5 };

struct TextArgs {
    const uint8_t *buf;
    uint64_t len, start;
    uint64_t first_tile, n_tiles; // tiles [first_tile, first_tile + n_tiles) cover [start, len)
    uint32_t comment, secondary;  // secondary >= 256: none
    // per tile (indexed from first_tile)
    uint32_t *t_rec, *t_hdr, *t_tabs_rec, *t_tabs_head;
    uint8_t *t_last_kind;
    uint64_t *t_last_ls;          // last line start in the tile, or ~0
    uint8_t *t_carry;             // kind of the line open at the tile's start
    uint64_t *rec_base, *hdr_base, *tab_base; // n_tiles + 1 each
    // per group of 1024 tiles: records, headers, record tabs whose line kind is
    // known in the group and those waiting on the kind carried in, the group's
    // last line kind and last line start (+1, 0 none); the scan turns the first
    // three into bases and sets the carried kind
    uint64_t *g_rec, *g_hdr, *g_known, *g_pending, *g_last_ls;
    uint32_t *g_last_kind, *g_carry;
    uint64_t *totals;             // TX_T_N
    // outputs
    uint64_t *record_offsets, *record_ends, *header_offsets, *tabs, *field_starts;
};

hipError_t launch_text_count(const TextArgs &a, hipStream_t s);
hipError_t launch_text_tiles(const TextArgs &a, hipStream_t s);
hipError_t launch_text_write(const TextArgs &a, hipStream_t s);

// Result words of a FASTA layout.
enum { FA_T_VALID = 0, FA_T_N = 1 };

struct FastaArgs {
    const uint8_t *buf;
    uint64_t len, n_tiles;
    uint32_t *t_gt, *t_nl;        // per tile: contig starts, newlines
    uint64_t *gt_base, *nl_base;  // n_tiles + 1 each
    uint64_t n_gt;                // contig starts found ('>' at a line start)
    uint64_t *starts;             // n_gt contig starts, in order
    uint64_t *seq_start;          // per contig: first sequence byte
    uint64_t *nonnl0;             // per contig: non-newline bytes before seq_start
    uint64_t *seq_len;            // per contig: bases (the SEQUENCE_LENGTHS column)
    uint64_t *seq_off;            // exclusive prefix of seq_len, n_gt + 1 entries
    uint64_t *result;             // FA_T_N words
    uint8_t *out;                 // the compacted sequence
};

hipError_t launch_fasta_count(const FastaArgs &a, hipStream_t s);
hipError_t launch_fasta_starts(const FastaArgs &a, hipStream_t s);
hipError_t launch_fasta_contigs(const FastaArgs &a, hipStream_t s);
hipError_t launch_fasta_compact(const FastaArgs &a, hipStream_t s);

} // namespace frhip
#endif
