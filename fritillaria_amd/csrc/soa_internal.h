// soa_internal.h -- device arrays and launchers of the BAM SoA extras.
#ifndef FRHIP_SOA_INTERNAL_H
#define FRHIP_SOA_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

// One aux field: tag, type byte, B subtype (0 otherwise), value offset from
// the record start, element count (1 for scalars; Z/H: string length without
// the NUL; B: array count) and value bytes (Z/H: with the NUL).
struct AuxEntry {
    uint8_t tag[2];
    uint8_t type;
    uint8_t subtype;
    uint32_t value_offset;
    uint32_t count;
    uint32_t size;
};
static_assert(sizeof(AuxEntry) == 16, "aux index entry");

// Records are taken in tiles of SOA_TILE (one per thread of a workgroup).
constexpr uint32_t SOA_TILE = 256;

// Result words of an SoA build: the three totals (CIGAR ops, packed sequence
// bytes, aux entries), then 1 when a total (or the record count) exceeds the
// capacity the outputs were sized for; the fill pass then writes nothing.
enum { SOA_TOT_OPS = 0, SOA_TOT_SEQ = 1, SOA_TOT_AUX = 2, SOA_OVERFLOW = 3, SOA_N_RESULT = 4 };
// aux_bounds of a record whose fields the fill pass walks again
constexpr uint32_t SOA_WALK = 0xFFFFFFFFu;

struct SoaArgs {
    const uint8_t *buf;
    uint64_t n;            // record count, or (with n_dev) the most records the grid covers
    const uint64_t *n_dev; // device-side record count (batched form), or null
    // decoded record columns: the record's place and its field bounds
    const uint64_t *record_offsets;
    const uint32_t *l_seq, *cigar_start, *seq_start, *aux_start, *record_end;
    // count pass, per record (cap_rec entries), and per tile; aux_bounds packs
    // the starts of aux fields 1..4 as byte offsets from aux_start (records of
    // at most 5 fields within 256 bytes), else SOA_WALK
    uint32_t *n_ops, *n_aux, *cg_at, *aux_bounds;
    uint64_t *tile_sums;   // 3 per tile: ops, sequence bytes, aux entries
    uint64_t *tile_base;   // 3 per tile: exclusive prefix of tile_sums
    uint64_t *result;      // SOA_N_RESULT words
    unsigned long long *first_bad; // lowest failing record (init ~0)
    // capacities of the per-record scratch and of the three data columns
    uint64_t cap_rec, cap_ops, cap_seq, cap_aux;
    // outputs: offsets (n + 1 each) and data
    uint64_t *op_off, *seq_off, *aux_off;
    uint32_t *ops;
    uint8_t *seq;
    AuxEntry *aux;
};

__host__ __device__ inline uint64_t soa_tiles(uint64_t n) { return (n + SOA_TILE - 1) / SOA_TILE; }

// Scratch of an SoA build over at most cap_rec records, carved at 256-byte
// alignment: the three per-record counts, the tile sums and bases, and the
// result words with the first failing record right after them.
struct SoaScratchLayout {
    uint64_t bytes, o_ops, o_aux, o_cg, o_bnd, o_sums, o_base, o_result;
};

inline SoaScratchLayout soa_scratch_layout(uint64_t cap_rec)
{
    SoaScratchLayout L{};
    uint64_t p = 0;
    auto take = [&](uint64_t b) { uint64_t o = p; p = (p + (b ? b : 1) + 255) & ~255ull; return o; };
    L.o_ops = take(cap_rec * 4);
    L.o_aux = take(cap_rec * 4);
    L.o_cg = take(cap_rec * 4);
    L.o_bnd = take(cap_rec * 4);
    L.o_sums = take(3 * soa_tiles(cap_rec) * 8);
    L.o_base = take(3 * soa_tiles(cap_rec) * 8);
    L.o_result = take((SOA_N_RESULT + 1) * 8);
    L.bytes = p;
    return L;
}

inline void soa_bind_scratch(SoaArgs &a, uint8_t *s, const SoaScratchLayout &L)
{
    a.n_ops = (uint32_t *)(s + L.o_ops);
    a.n_aux = (uint32_t *)(s + L.o_aux);
    a.cg_at = (uint32_t *)(s + L.o_cg);
    a.aux_bounds = (uint32_t *)(s + L.o_bnd);
    a.tile_sums = (uint64_t *)(s + L.o_sums);
    a.tile_base = (uint64_t *)(s + L.o_base);
    a.result = (uint64_t *)(s + L.o_result);
    a.first_bad = (unsigned long long *)(a.result + SOA_N_RESULT);
}

inline void soa_bind_columns(SoaArgs &a, const uint64_t *record_offsets, const uint32_t *l_seq,
                             const uint32_t *cigar_start, const uint32_t *seq_start, const uint32_t *aux_start,
                             const uint32_t *record_end)
{
    a.record_offsets = record_offsets;
    a.l_seq = l_seq;
    a.cigar_start = cigar_start;
    a.seq_start = seq_start;
    a.aux_start = aux_start;
    a.record_end = record_end;
}

// Count pass: per-record op / aux counts and the long-CIGAR source, per-tile
// sums, the first malformed record.
hipError_t launch_soa_count(const SoaArgs &a, hipStream_t s);
// One workgroup: tile_base from tile_sums, the totals and the overflow flag.
hipError_t launch_soa_scan(const SoaArgs &a, hipStream_t s);
// Fill pass: the three offset columns and the ops, packed sequence and aux
// entries, each record's data at its offsets.
hipError_t launch_soa_fill(const SoaArgs &a, hipStream_t s);

} // namespace frhip
#endif
