// frhip_internal.h -- kernel argument blocks and launchers shared by the HIP
// sources and the host implementation. Not part of the public ABI.
#ifndef FRHIP_INTERNAL_H
#define FRHIP_INTERNAL_H
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace frhip {

struct InflateArgs {
    const uint8_t *in;       // compressed bytes (padded by at least 4 KiB)
    const uint64_t *in_off;  // payload start per member, absolute in `in`
    const uint32_t *in_len;  // payload length per member
    uint8_t *out;            // inflated output
    const uint64_t *out_off; // output offset per member (prefix sum of ISIZE)
    const uint32_t *out_cap; // ISIZE per member
    uint32_t *status;
    uint32_t *produced;
    uint32_t *crc;
    uint32_t n;
    uint32_t *queue;         // work-queue counter (persistent form) or null
    const uint32_t *order;   // member order for the queue, or null
    const uint32_t *x2n;     // x^(2^k) mod P, k = 0..31
    uint32_t *nrec;          // two-phase schedules: match records per member
    uint32_t *rec;           // two-phase schedules: match record lists
    uint32_t two_phase_min;  // two-phase schedules: fewer members run the one-wave form
};

// Inflate kernel variants: output ring bytes per member and waves per workgroup.
struct InflateSchedule {
    uint32_t ring;
    uint32_t waves_per_group;
    uint32_t tokens; // 0 scalar decode into the ring; 1 tokens, resolved 64 at a time; 2 lane-speculative
                     // tokens; 3 two-phase (lane-per-member decode, then wave-per-member copies)
};
extern const InflateSchedule INFLATE_SCHEDULES[];
extern const int N_INFLATE_SCHEDULES;
constexpr int DEFAULT_INFLATE_SCHEDULE = 16;

// Launches `waves` waves (one member each at a time) of schedule `schedule`.
hipError_t launch_inflate(const InflateArgs &a, uint32_t waves, hipStream_t stream, int schedule);
// Device scratch a schedule needs for n members of out_bytes inflated bytes in
// total (0 for all but the two-phase schedules), and how it is laid out.
uint64_t inflate_scratch_bytes(int schedule, uint64_t out_bytes, uint64_t n);
void inflate_scratch_bind(InflateArgs &a, void *scratch, uint64_t n);
uint32_t inflate_lds_bytes(int schedule);
uint32_t inflate_waves_per_cu(int schedule);

// Record-boundary discovery for BAM over an inflated batch.
constexpr uint32_t BAM_K = 8;                 // survivors kept per block
constexpr uint64_t BAM_NONE = ~0ull;          // block has no entry
// most records that can start in one member (BCF's are >= 32 B), rounded up
// to a multiple of 8 so that each member's u16 offsets start 16-byte aligned
constexpr uint32_t BAM_CAPB = (65536 / 32 + 1 + 7) & ~7u;
// the walk stores offsets eight at a time as one 16-byte store into each
// member's slab: the slab must hold a member's records and stay 16-byte aligned
static_assert(BAM_CAPB % 8 == 0, "member slabs of u16 offsets must be a multiple of 16 bytes");
static_assert(BAM_CAPB >= 65536 / 32 + 1, "a member slab must hold every record a member can start");
constexpr uint32_t BAM_MODE_FALLBACK = 255;

// Length-prefixed record chains the boundary scan handles.
enum : uint32_t { CHAIN_BAM = 0, CHAIN_BCF = 1 };

struct BamSpecArgs {
    uint32_t format;           // CHAIN_*
    uint32_t samples, contigs; // BCF: the header's sample and contig counts (candidate filter)
    const uint8_t *buf;
    uint64_t len;
    const uint64_t *block_starts; // nb + 1
    uint32_t nb;
    uint32_t window;
    uint8_t *spec_ok;
    uint8_t *n_surv;
    uint64_t *surv_pos;  // nb * BAM_K
    uint32_t *surv_cnt;  // nb * BAM_K
    uint64_t *land;
    uint64_t *exitp;
    uint32_t *cnt_tail;
    uint8_t *complete;
    uint16_t *rel_offs;  // nb * BAM_CAPB
    // Several independent buffers in one launch (queries of a serving batch):
    // per member, where its walks stop and where its buffer ends. Null: the
    // member end and `len`.
    const uint64_t *mem_until;
    const uint64_t *mem_len;
};

struct BamReconcileArgs {
    uint32_t format;
    const uint8_t *buf;
    uint64_t len;
    const uint64_t *block_starts;
    uint32_t nb;
    uint64_t start;
    const uint64_t *start_dev; // device-side start (batched form), or null
    const uint64_t *len_dev;   // device-side buffer length (batched form), or null
    const uint8_t *spec_ok;
    const uint8_t *n_surv;
    const uint64_t *surv_pos;
    const uint32_t *surv_cnt;
    const uint64_t *exitp;
    const uint32_t *cnt_tail;
    const uint8_t *complete;
    uint64_t *entry;
    uint32_t *count;
    uint64_t *first_index;
    uint8_t *mode;
    uint64_t *totals;    // {records, tail, status, fallbacks}
    // Segments (one workgroup each): members [seg_mb[q], seg_mb[q+1]) form one
    // chain from seg_start[q] over a buffer ending at seg_len[q]; totals go to
    // seg_totals[4q..4q+3] and first_index counts from the segment's first
    // record. Null seg_mb: one segment of all nb members, start/len/totals above.
    const uint32_t *seg_mb;
    const uint64_t *seg_start;
    const uint64_t *seg_len;
    uint64_t *seg_totals;
    const uint64_t *mem_until; // as in BamSpecArgs
    // Single-chain form only: scratch for the parallel fast path ((nb + 1)
    // counts scanned, 4 words of state), or null for the serial reconcile alone.
    uint64_t *fast_scan;
    uint32_t *fast_state;
};

struct BamEmitArgs {
    uint32_t format;
    const uint8_t *buf;
    const uint64_t *block_starts;
    const uint64_t *coffsets; // compressed offset per block (for virtual offsets), or null
    uint32_t nb;
    const uint64_t *entry;
    const uint32_t *count;
    const uint64_t *first_index;
    const uint8_t *mode;
    const uint32_t *surv_cnt;
    const uint16_t *rel_offs;
    uint64_t *record_offsets;
    uint64_t *virtual_offsets; // or null
    const uint32_t *mem_seg;   // segment of each member, or null (one segment)
    const uint64_t *seg_base;  // first record index of each segment, or null
};

struct BamColumns {
    int32_t *ref_id;
    int32_t *pos;
    uint16_t *flags;
    uint8_t *mapq;
    uint32_t *l_seq;
    int32_t *next_ref_id;
    int32_t *next_pos;
    int32_t *tlen;
    uint64_t *record_offsets;
    uint32_t *cigar_start;
    uint32_t *seq_start;
    uint32_t *qual_start;
    uint32_t *aux_start;
    uint32_t *record_end;
};

struct BamDecodeArgs {
    const uint8_t *buf;
    uint64_t len; // bytes of buf that may be read (records lie inside it)
    uint64_t n;
    const uint64_t *n_dev; // device-side record count (batched form), or null
    BamColumns c;
};

// BCF site-core columns and typed-value bounds
// (crates/fritillaria-bcf/src/columnar/device.rs:59-98 order).
struct BcfColumns {
    int32_t *chrom, *pos, *rlen;
    uint32_t *qual;
    uint16_t *n_info, *n_allele;
    uint32_t *n_sample;
    uint8_t *n_fmt;
    uint64_t *record_offsets;
    uint32_t *alleles_start, *filters_start, *info_start, *genotypes_start, *record_end;
};

struct BcfDecodeArgs {
    const uint8_t *buf;
    uint64_t n;
    uint64_t tail;           // end of the last record
    BcfColumns c;
    unsigned long long *first_bad; // lowest record index whose typed values do not parse (init ~0)
};

hipError_t launch_bcf_decode(const BcfDecodeArgs &a, hipStream_t s);
hipError_t launch_bam_spec(const BamSpecArgs &a, hipStream_t s);
hipError_t launch_bam_reconcile(const BamReconcileArgs &a, hipStream_t s);
hipError_t launch_bam_emit(const BamEmitArgs &a, hipStream_t s);
hipError_t launch_bam_decode(const BamDecodeArgs &a, uint64_t max_records, hipStream_t s);
hipError_t launch_bam_reconcile_segments(const BamReconcileArgs &a, uint32_t n_seg, hipStream_t s);
// seg_base[q] = sum of seg_totals[4p] for p < q; seg_base[n_seg] = the total
hipError_t launch_bam_segment_bases(const uint64_t *seg_totals, uint32_t n_seg, uint64_t *seg_base, hipStream_t s);

} // namespace frhip
#endif
