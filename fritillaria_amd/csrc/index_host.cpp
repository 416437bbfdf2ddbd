// index_host.cpp -- BAI / CSI binning indexes and the region -> chunk mapping.
//
// An index is small (kilobytes to megabytes) and a query touches a handful of
// bins, so this is host code: it turns a region into the merged chunk list
// that frhip_server_launch_regions reads on the device. The layouts and rules
// follow the reference's readers and BinningIndex::query:
//   BAI  crates/fritillaria-bam/src/bai/io/reader/index.rs:9-37,
//        .../index/reference_sequences.rs:13-33, .../reference_sequences/bins.rs:8-58,
//        .../reference_sequences/intervals.rs:7-22
//   CSI  crates/fritillaria-csi/src/io/reader/index.rs:72-120, .../reference_sequences.rs:50-86,
//        .../reference_sequences/bins.rs:69-108, .../reference_sequences/metadata.rs (two chunks)
//   query crates/fritillaria-csi/src/binning_index/index.rs:93-167 with reg2bins
//        (.../index/reference_sequence.rs:376-397), optimize_chunks
//        (crates/fritillaria-csi/src/binning_index.rs:137-160) and min_offset
//        (.../index/reference_sequence/index/linear_index.rs:11-15, binned_index.rs:12-28).
#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/frhip.h"
#include "../../include/frhip_index.h"
#include "host_internal.h"

using namespace frhip;

namespace {

struct Bin {
    uint32_t id;
    uint64_t loff; // CSI: the bin's linear offset; BAI: unused
    std::vector<frhip_chunk> chunks;
};

struct RefIx {
    std::vector<Bin> bins; // sorted by id
    std::vector<uint64_t> linear;
    bool has_meta = false;
    uint64_t meta[4] = {};
};

struct Cursor {
    const uint8_t *p;
    uint64_t n, at = 0;
    bool ok = true;
    bool need(uint64_t k)
    {
        if (!ok || n - at < k) ok = false;
        return ok;
    }
    uint32_t u32()
    {
        if (!need(4)) return 0;
        uint32_t v;
        std::memcpy(&v, p + at, 4);
        at += 4;
        return v;
    }
    int32_t i32() { return (int32_t)u32(); }
    uint64_t u64()
    {
        if (!need(8)) return 0;
        uint64_t v;
        std::memcpy(&v, p + at, 8);
        at += 8;
        return v;
    }
};

uint32_t bin_limit(uint32_t depth) { return (uint32_t)((1ull << ((depth + 1) * 3)) / 7); }

// [beg, end] 0-based closed; CSIv1 reg2bin
uint32_t reg2bin(uint64_t beg, uint64_t end, uint32_t min_shift, uint32_t depth)
{
    uint32_t s = min_shift;
    uint64_t t = ((1ull << (depth * 3)) - 1) / 7;
    for (uint32_t l = depth; l > 0; --l) {
        if (beg >> s == end >> s) return (uint32_t)(t + (beg >> s));
        s += 3;
        t -= 1ull << ((l - 1) * 3);
    }
    return 0;
}

} // namespace

struct frhip_index {
    int format = 0;
    uint32_t min_shift = 14, depth = 5;
    std::vector<RefIx> refs;
    bool has_unplaced = false;
    uint64_t unplaced = 0;

    const Bin *find(const RefIx &r, uint64_t id) const
    {
        auto it = std::lower_bound(r.bins.begin(), r.bins.end(), id,
                                   [](const Bin &b, uint64_t v) { return b.id < v; });
        return it != r.bins.end() && it->id == id ? &*it : nullptr;
    }
};

namespace {

int bad(frhip_error *err, uint64_t at, const char *what)
{
    return set_err(err, FRHIP_E_IO, at, "invalid index: %s", what);
}

int read_chunks(Cursor &c, std::vector<frhip_chunk> &out, frhip_error *err)
{
    const int32_t n = c.i32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    if (n < 0) return bad(err, c.at, "invalid chunk count");
    if (!c.need((uint64_t)n * 16)) return bad(err, c.at, "unexpected end");
    out.resize((size_t)n);
    for (int32_t k = 0; k < n; ++k) {
        out[k].start = c.u64();
        out[k].end = c.u64();
    }
    return FRHIP_OK;
}

int read_meta(Cursor &c, RefIx &r, frhip_error *err)
{
    const uint32_t n = c.u32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    if (n != 2) return bad(err, c.at, "invalid metadata chunk count");
    if (r.has_meta) return bad(err, c.at, "duplicate metadata bin");
    for (auto &m : r.meta) m = c.u64();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    r.has_meta = true;
    return FRHIP_OK;
}

int finish_bins(RefIx &r, uint64_t at, frhip_error *err)
{
    std::sort(r.bins.begin(), r.bins.end(), [](const Bin &a, const Bin &b) { return a.id < b.id; });
    for (size_t k = 1; k < r.bins.size(); ++k)
        if (r.bins[k].id == r.bins[k - 1].id) return bad(err, at, "duplicate bin ID");
    return FRHIP_OK;
}

int read_bai(Cursor &c, frhip_index *ix, frhip_error *err)
{
    ix->format = FRHIP_INDEX_BAI;
    ix->min_shift = 14;
    ix->depth = 5;
    const uint32_t meta_id = bin_limit(ix->depth) + 1;
    const uint32_t n_ref = c.u32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    // every reference takes at least 8 bytes (n_bin, n_intv)
    if (!c.need((uint64_t)n_ref * 8)) return bad(err, c.at, "unexpected end");
    ix->refs.resize(n_ref);
    for (RefIx &r : ix->refs) {
        const uint32_t n_bin = c.u32();
        if (!c.ok || !c.need((uint64_t)n_bin * 8)) return bad(err, c.at, "unexpected end");
        r.bins.reserve(n_bin);
        for (uint32_t b = 0; b < n_bin; ++b) {
            const uint32_t id = c.u32();
            if (!c.ok) return bad(err, c.at, "unexpected end");
            int rc;
            if (id == meta_id) {
                rc = read_meta(c, r, err);
            } else {
                r.bins.push_back(Bin{id, 0, {}});
                rc = read_chunks(c, r.bins.back().chunks, err);
            }
            if (rc != FRHIP_OK) return rc;
        }
        if (int rc = finish_bins(r, c.at, err)) return rc;
        const uint32_t n_intv = c.u32();
        if (!c.ok || !c.need((uint64_t)n_intv * 8)) return bad(err, c.at, "unexpected end");
        r.linear.resize(n_intv);
        for (auto &v : r.linear) v = c.u64();
    }
    return FRHIP_OK;
}

int read_csi(Cursor &c, frhip_index *ix, frhip_error *err)
{
    ix->format = FRHIP_INDEX_CSI;
    const int32_t ms = c.i32(), dp = c.i32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    if (ms < 0 || ms > 255) return bad(err, c.at, "invalid min shift");
    if (dp < 0 || dp > 255) return bad(err, c.at, "invalid depth");
    // the binning must address positions in 63 bits
    if (ms + 3 * dp > 62 || dp > 10) return bad(err, c.at, "unsupported min shift and depth");
    ix->min_shift = (uint32_t)ms;
    ix->depth = (uint32_t)dp;
    const int32_t l_aux = c.i32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    if (l_aux < 0) return bad(err, c.at, "invalid aux length");
    if (!c.need((uint64_t)l_aux)) return bad(err, c.at, "unexpected end");
    c.at += (uint64_t)l_aux;
    const uint32_t meta_id = bin_limit(ix->depth) + 1;
    const int32_t n_ref = c.i32();
    if (!c.ok) return bad(err, c.at, "unexpected end");
    if (n_ref < 0) return bad(err, c.at, "invalid reference sequence count");
    if (!c.need((uint64_t)n_ref * 4)) return bad(err, c.at, "unexpected end");
    ix->refs.resize((size_t)n_ref);
    for (RefIx &r : ix->refs) {
        const int32_t n_bin = c.i32();
        if (!c.ok) return bad(err, c.at, "unexpected end");
        if (n_bin < 0) return bad(err, c.at, "invalid bin count");
        if (!c.need((uint64_t)n_bin * 16)) return bad(err, c.at, "unexpected end");
        r.bins.reserve((size_t)n_bin);
        for (int32_t b = 0; b < n_bin; ++b) {
            const uint32_t id = c.u32();
            const uint64_t loff = c.u64();
            if (!c.ok) return bad(err, c.at, "unexpected end");
            int rc;
            if (id == meta_id) {
                rc = read_meta(c, r, err);
            } else {
                r.bins.push_back(Bin{id, loff, {}});
                rc = read_chunks(c, r.bins.back().chunks, err);
            }
            if (rc != FRHIP_OK) return rc;
        }
        if (int rc = finish_bins(r, c.at, err)) return rc;
    }
    return FRHIP_OK;
}

} // namespace

extern "C" int frhip_index_read(const uint8_t *data, uint64_t len, frhip_index **out, frhip_error *err)
{
    *out = nullptr;
    if (!data || len < 4) return bad(err, 0, "unexpected end");
    frhip_index *ix = new (std::nothrow) frhip_index;
    if (!ix) return set_err(err, FRHIP_E_IO, 0, "host allocation failed");
    Cursor c{data, len};
    c.at = 4;
    int rc;
    try {
        if (std::memcmp(data, "BAI\1", 4) == 0) rc = read_bai(c, ix, err);
        else if (std::memcmp(data, "CSI\1", 4) == 0) rc = read_csi(c, ix, err);
        else rc = bad(err, 0, "neither a BAI nor a CSI magic number");
    } catch (const std::bad_alloc &) {
        rc = set_err(err, FRHIP_E_IO, c.at, "host allocation failed");
    }
    if (rc == FRHIP_OK && c.at != len) {
        // the optional unplaced, unmapped record count
        if (len - c.at < 8) rc = bad(err, c.at, "truncated unplaced unmapped record count");
        else {
            ix->has_unplaced = true;
            ix->unplaced = c.u64();
        }
    }
    if (rc != FRHIP_OK) {
        delete ix;
        return rc;
    }
    *out = ix;
    return FRHIP_OK;
}

extern "C" void frhip_index_free(frhip_index *ix) { delete ix; }

extern "C" int frhip_index_format(const frhip_index *ix, uint32_t *min_shift, uint32_t *depth)
{
    if (min_shift) *min_shift = ix->min_shift;
    if (depth) *depth = ix->depth;
    return ix->format;
}

extern "C" uint32_t frhip_index_references(const frhip_index *ix) { return (uint32_t)ix->refs.size(); }

extern "C" int frhip_index_unplaced_unmapped(const frhip_index *ix, uint64_t *n)
{
    if (!ix->has_unplaced) return 0;
    if (n) *n = ix->unplaced;
    return 1;
}

extern "C" int frhip_index_metadata(const frhip_index *ix, uint32_t ref_id, uint64_t out[4])
{
    if (ref_id >= ix->refs.size() || !ix->refs[ref_id].has_meta) return 0;
    std::memcpy(out, ix->refs[ref_id].meta, sizeof ix->refs[ref_id].meta);
    return 1;
}

extern "C" int frhip_index_last_first_record_start(const frhip_index *ix, uint64_t *vpos)
{
    for (size_t k = ix->refs.size(); k-- > 0;) {
        const RefIx &r = ix->refs[k];
        if (ix->format == FRHIP_INDEX_BAI) {
            if (r.linear.empty()) continue;
            *vpos = r.linear.back();
            return 1;
        }
        if (r.bins.empty()) continue;
        uint64_t m = 0;
        for (const Bin &b : r.bins) m = std::max(m, b.loff);
        *vpos = m;
        return 1;
    }
    return 0;
}

extern "C" int frhip_index_query(const frhip_index *ix, uint32_t ref_id, uint64_t start, uint64_t end,
                                 frhip_chunk *out, uint64_t cap, uint64_t *n, frhip_error *err)
{
    *n = 0;
    if (ref_id >= ix->refs.size())
        return set_err(err, FRHIP_E_IO, 0, "invalid reference sequence ID: %u", ref_id);
    const RefIx &r = ix->refs[ref_id];
    const uint32_t ms = ix->min_shift, depth = ix->depth;
    const uint64_t max_pos = (1ull << (ms + 3 * depth)) - 1;
    const uint64_t s = start ? start : 1;
    if (s > max_pos) return set_err(err, FRHIP_E_IO, 0, "invalid start bound");
    const uint64_t e = end ? end : max_pos;
    if (e > max_pos) return set_err(err, FRHIP_E_IO, 0, "invalid end bound");
    // chunks of every bin reg2bins names: at level l the bins t + (beg >> sh) ..= t + (end >> sh)
    const uint64_t beg0 = s - 1, end0 = e - 1;
    std::vector<frhip_chunk> cs;
    uint64_t t = 0;
    for (uint32_t l = 0; l <= depth; ++l) {
        const uint32_t sh = ms + 3 * (depth - l);
        const uint64_t lo = t + (beg0 >> sh), hi = t + (end0 >> sh);
        auto it = std::lower_bound(r.bins.begin(), r.bins.end(), lo,
                                   [](const Bin &b, uint64_t v) { return b.id < v; });
        for (; it != r.bins.end() && it->id <= hi; ++it) cs.insert(cs.end(), it->chunks.begin(), it->chunks.end());
        t += 1ull << (l * 3);
    }
    // min_offset: the linear entry of the start's window (BAI), or the
    // loffset of the start's bin or its nearest ancestor present (CSI)
    uint64_t min_off = 0;
    if (ix->format == FRHIP_INDEX_BAI) {
        const uint64_t w = (s - 1) >> ms;
        if (w < r.linear.size()) min_off = r.linear[w];
    } else {
        uint64_t id = reg2bin(s - 1, s - 1, ms, depth);
        for (;;) {
            if (const Bin *b = ix->find(r, id)) {
                min_off = b->loff;
                break;
            }
            if (id == 0) break;
            id = (id - 1) / 8;
        }
    }
    // optimize_chunks: drop chunks ending at or before min_off, sort, merge
    // overlapping and touching chunks
    std::vector<frhip_chunk> kept;
    kept.reserve(cs.size());
    for (const frhip_chunk &c : cs)
        if (c.end > min_off) kept.push_back(c);
    std::sort(kept.begin(), kept.end(), [](const frhip_chunk &a, const frhip_chunk &b) { return a.start < b.start; });
    std::vector<frhip_chunk> merged;
    for (const frhip_chunk &c : kept) {
        if (merged.empty() || c.start > merged.back().end) merged.push_back(c);
        else if (merged.back().end < c.end) merged.back().end = c.end;
    }
    *n = merged.size();
    const uint64_t k = std::min<uint64_t>(cap, merged.size());
    if (k && out) std::memcpy(out, merged.data(), k * sizeof(frhip_chunk));
    return FRHIP_OK;
}
