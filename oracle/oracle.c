/*
 * oracle.c -- CPU restatement of fritillaria's device path, used ONLY as a
 * test checker (pytest, __graft_entry__.smoke, bench.py's cpu_baseline leg).
 * It is never linked into, or called by, the product library.
 *
 * Every function cites the reference file:line whose behaviour it restates.
 * Parity pins (what checks each function against something independent):
 *   or_crc32            zlib.crc32 on random data; reference known answers
 *                       crc32("noodles")=0x802a58a1, crc32("")=0
 *                       (crates/fritillaria-bgzf/src/deflate.rs:116-119).
 *   or_inflate_member   zlib raw inflate (windowBits -15) on stored / fixed /
 *                       dynamic streams at every level; hand-built corrupt
 *                       streams for every status code; fixture trailers.
 *   or_parse_header,
 *   or_discover         EOF block bytes, BC+1 rule, fixture member tables
 *                       (crates/fritillaria-bgzf/src/block.rs:40-43,192-215).
 *   or_bam_*            samtools-derived known answers of htslib.bam,
 *                       htslib_multiblock.bam and the ONT size table
 *                       (crates/fritillaria-bam/tests/htslib.rs:91-186,284-316;
 *                       testdata/README.md:147-152); brute-force equality of the
 *                       blocked scan with the plain chain walk over every
 *                       blocking (crates/fritillaria-bam/src/columnar/blocked.rs:306-319).
 *   or_virtual_offset   noodles vectors (crates/fritillaria-bgzf/src/io/block.rs:61-77).
 *   or_fastq_*          fixture record counts and read-length ranges
 *                       (testdata/README.md:319-322); the FASTQ fixtures are
 *                       `samtools fastq` of the BAM fixtures, so each record's
 *                       name, bases and qualities must equal the BAM record's
 *                       (reverse-complemented for flag 0x10), decoded by the
 *                       independent BAM oracle; the reference's hand-built
 *                       decoy, seam and short-quality cases
 *                       (crates/fritillaria-fastq/src/columnar/record.rs:291-468).
 */
#include "oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* CRC-32/ISO-HDLC, bitwise (poly 0xEDB88320 reflected, init/xorout ~0).      */
/* Restates crates/fritillaria-cuda/kernels/inflate.cu:387-397.               */
/* ------------------------------------------------------------------------ */
uint32_t or_crc32(const uint8_t *p, uint64_t len)
{
    uint32_t crc = 0xFFFFFFFFu;
    for (uint64_t i = 0; i < len; ++i) {
        crc ^= p[i];
        for (int k = 0; k < 8; ++k)
            crc = (crc >> 1) ^ (0xEDB88320u & (uint32_t)(-(int32_t)(crc & 1u)));
    }
    return ~crc;
}

/* ------------------------------------------------------------------------ */
/* RFC 1951 inflate, bit-serial canonical decode.                            */
/* Restates crates/fritillaria-cuda/kernels/inflate.cu:23-397 (puff-shaped),  */
/* including its status taxonomy and the order in which errors are detected. */
/* ------------------------------------------------------------------------ */
#define MAXBITS 15
#define MAXLCODES 286
#define MAXDCODES 30
#define FIXLCODES 288

static const short LEN_BASE[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                   35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
static const short LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                    2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
static const int DIST_BASE[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129,
                                  193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097,
                                  6145, 8193, 12289, 16385, 24577};
static const short DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                     6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
static const short CLEN_ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

typedef struct {
    const uint8_t *in;
    uint32_t in_len, in_pos;
    uint32_t bitbuf;
    int bitcnt;
    uint8_t *out;
    uint32_t out_cap, out_pos;
    int err;
} st_t;

static int bits(st_t *s, int need)
{
    uint64_t val = s->bitbuf;
    while (s->bitcnt < need) {
        if (s->in_pos >= s->in_len) {
            s->err = OR_ERR_TRUNCATED;
            return 0;
        }
        val |= (uint64_t)s->in[s->in_pos++] << s->bitcnt;
        s->bitcnt += 8;
    }
    s->bitbuf = (uint32_t)(val >> need);
    s->bitcnt -= need;
    return (int)(val & ((1ull << need) - 1));
}

typedef struct {
    short count[MAXBITS + 1];
    short symbol[FIXLCODES];
} huff_t;

static int decode(st_t *s, const huff_t *h)
{
    int code = 0, first = 0, index = 0;
    for (int len = 1; len <= MAXBITS; ++len) {
        code |= bits(s, 1);
        if (s->err) return -1;
        int count = h->count[len];
        if (code - first < count) return h->symbol[index + (code - first)];
        index += count;
        first = (first + count) << 1;
        code <<= 1;
    }
    return -1;
}

/* 0 complete, >0 incomplete, <0 over-subscribed (inflate.cu:126-161). */
static int construct(huff_t *h, const short *length, int n)
{
    for (int len = 0; len <= MAXBITS; ++len) h->count[len] = 0;
    for (int sym = 0; sym < n; ++sym) h->count[length[sym]]++;
    if (h->count[0] == n) return 0;
    int left = 1;
    for (int len = 1; len <= MAXBITS; ++len) {
        left <<= 1;
        left -= h->count[len];
        if (left < 0) return left;
    }
    short offs[MAXBITS + 1];
    offs[1] = 0;
    for (int len = 1; len < MAXBITS; ++len) offs[len + 1] = offs[len] + h->count[len];
    for (int sym = 0; sym < n; ++sym)
        if (length[sym] != 0) h->symbol[offs[length[sym]]++] = (short)sym;
    return left;
}

static void codes(st_t *s, const huff_t *lc, const huff_t *dc)
{
    for (;;) {
        int sym = decode(s, lc);
        if (s->err) return;
        if (sym < 0) { s->err = OR_ERR_BAD_CODE; return; }
        if (sym < 256) {
            if (s->out_pos >= s->out_cap) { s->err = OR_ERR_OUTPUT_OVERFLOW; return; }
            s->out[s->out_pos++] = (uint8_t)sym;
            continue;
        }
        if (sym == 256) return;
        sym -= 257;
        if (sym >= 29) { s->err = OR_ERR_BAD_CODE; return; }
        int len = LEN_BASE[sym] + bits(s, LEN_EXTRA[sym]);
        if (s->err) return;
        sym = decode(s, dc);
        if (s->err) return;
        if (sym < 0 || sym >= MAXDCODES) { s->err = OR_ERR_BAD_DISTANCE; return; }
        int dist = DIST_BASE[sym] + bits(s, DIST_EXTRA[sym]);
        if (s->err) return;
        if ((uint32_t)dist > s->out_pos) { s->err = OR_ERR_BAD_DISTANCE; return; }
        if (s->out_pos + (uint32_t)len > s->out_cap) { s->err = OR_ERR_OUTPUT_OVERFLOW; return; }
        for (int i = 0; i < len; ++i) {
            s->out[s->out_pos] = s->out[s->out_pos - dist];
            s->out_pos++;
        }
    }
}

static void stored(st_t *s)
{
    s->bitbuf = 0;
    s->bitcnt = 0;
    if (s->in_pos + 4 > s->in_len) { s->err = OR_ERR_TRUNCATED; return; }
    uint32_t len = s->in[s->in_pos] | ((uint32_t)s->in[s->in_pos + 1] << 8);
    uint32_t nlen = s->in[s->in_pos + 2] | ((uint32_t)s->in[s->in_pos + 3] << 8);
    s->in_pos += 4;
    if (len != (~nlen & 0xffffu)) { s->err = OR_ERR_STORED_LEN; return; }
    if (s->in_pos + len > s->in_len) { s->err = OR_ERR_TRUNCATED; return; }
    if (s->out_pos + len > s->out_cap) { s->err = OR_ERR_OUTPUT_OVERFLOW; return; }
    memcpy(s->out + s->out_pos, s->in + s->in_pos, len);
    s->out_pos += len;
    s->in_pos += len;
}

static void fixed(st_t *s)
{
    static huff_t lc, dc;
    static int built = 0;
    if (!built) {
        short lengths[FIXLCODES];
        int sym = 0;
        for (; sym < 144; ++sym) lengths[sym] = 8;
        for (; sym < 256; ++sym) lengths[sym] = 9;
        for (; sym < 280; ++sym) lengths[sym] = 7;
        for (; sym < FIXLCODES; ++sym) lengths[sym] = 8;
        construct(&lc, lengths, FIXLCODES);
        for (sym = 0; sym < 32; ++sym) lengths[sym] = 5;
        construct(&dc, lengths, 32);
        built = 1;
    }
    codes(s, &lc, &dc);
}

static void dynamic(st_t *s)
{
    short lengths[MAXLCODES + MAXDCODES];
    huff_t lc, dc;
    int nlen = bits(s, 5) + 257;
    int ndist = bits(s, 5) + 1;
    int ncode = bits(s, 4) + 4;
    if (s->err) return;
    if (nlen > MAXLCODES || ndist > MAXDCODES) { s->err = OR_ERR_BAD_COUNTS; return; }
    for (int i = 0; i < 19; ++i) lengths[CLEN_ORDER[i]] = 0;
    for (int i = 0; i < ncode; ++i) {
        lengths[CLEN_ORDER[i]] = (short)bits(s, 3);
        if (s->err) return;
    }
    if (construct(&lc, lengths, 19) != 0) { s->err = OR_ERR_BAD_TABLE; return; }
    int index = 0;
    while (index < nlen + ndist) {
        int sym = decode(s, &lc);
        if (s->err) return;
        if (sym < 0) { s->err = OR_ERR_BAD_CODE; return; }
        if (sym < 16) { lengths[index++] = (short)sym; continue; }
        short len = 0;
        if (sym == 16) {
            if (index == 0) { s->err = OR_ERR_BAD_LENGTH_REPEAT; return; }
            len = lengths[index - 1];
            sym = 3 + bits(s, 2);
        } else if (sym == 17) {
            sym = 3 + bits(s, 3);
        } else {
            sym = 11 + bits(s, 7);
        }
        if (s->err) return;
        if (index + sym > nlen + ndist) { s->err = OR_ERR_BAD_LENGTH_REPEAT; return; }
        while (sym-- > 0) lengths[index++] = len;
    }
    if (lengths[256] == 0) { s->err = OR_ERR_BAD_TABLE; return; }
    int e = construct(&lc, lengths, nlen);
    if (e != 0 && (e < 0 || nlen != lc.count[0] + 1)) { s->err = OR_ERR_BAD_TABLE; return; }
    e = construct(&dc, lengths + nlen, ndist);
    if (e != 0 && (e < 0 || ndist != dc.count[0] + 1)) { s->err = OR_ERR_BAD_TABLE; return; }
    codes(s, &lc, &dc);
}

/* inflate.cu:399-458 (one member): status, bytes produced. */
int or_inflate_member(const uint8_t *in, uint32_t in_len, uint8_t *out, uint32_t out_cap,
                      uint32_t *produced)
{
    st_t s;
    memset(&s, 0, sizeof s);
    s.in = in;
    s.in_len = in_len;
    s.out = out;
    s.out_cap = out_cap;
    int last;
    do {
        last = bits(&s, 1);
        int type = bits(&s, 2);
        if (s.err) break;
        if (type == 0) stored(&s);
        else if (type == 1) fixed(&s);
        else if (type == 2) dynamic(&s);
        else s.err = OR_ERR_BLOCK_TYPE;
    } while (s.err == 0 && !last);
    *produced = s.out_pos;
    return s.err;
}

/* ------------------------------------------------------------------------ */
/* BGZF framing. crates/fritillaria-bgzf/src/block.rs:88-164 (lenient: BC    */
/* anywhere among the subfields) and discover.rs:51-113 (the member walk).   */
/* ------------------------------------------------------------------------ */
static uint16_t rd16(const uint8_t *p) { return (uint16_t)(p[0] | (p[1] << 8)); }
static uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

int or_parse_header(const uint8_t *buf, uint64_t len, uint32_t *block_size, uint32_t *header_size)
{
    if (len < 12) return OR_BLK_TRUNCATED_HEADER;
    if (buf[0] != 0x1f || buf[1] != 0x8b) return OR_BLK_BAD_MAGIC;
    if (buf[2] != 8) return OR_BLK_BAD_METHOD;
    if ((buf[3] & 0x04) == 0) return OR_BLK_NO_FEXTRA;
    uint32_t xlen = rd16(buf + 10);
    uint32_t hs = 12 + xlen;
    if (len < hs) return OR_BLK_TRUNCATED_EXTRA;
    /* find_bc_subfield (block.rs:147-164) */
    const uint8_t *x = buf + 12;
    uint32_t xl = xlen;
    int found = 0;
    uint32_t bsize = 0;
    while (xl >= 4) {
        uint32_t slen = rd16(x + 2);
        if (4 + slen > xl) break;
        if (x[0] == 'B' && x[1] == 'C') {
            if (slen != 2) break;
            bsize = rd16(x + 4);
            found = 1;
            break;
        }
        x += 4 + slen;
        xl -= 4 + slen;
    }
    if (!found) return OR_BLK_NO_BC;
    uint32_t bs = bsize + 1;
    if (bs > 65536) return OR_BLK_TOO_BIG; /* unreachable with u16+1, kept for the order */
    if (bs < hs + 8) return OR_BLK_TOO_SMALL;
    *block_size = bs;
    *header_size = hs;
    return 0;
}

/*
 * discover.rs:51-113. Writes up to `cap` spans; returns the number written.
 * *position = resume point; *err = 0 or an OR_BLK_* code (then *err_offset is
 * the absolute offset of the failing member).
 */
uint64_t or_discover(const uint8_t *buf, uint64_t len, uint64_t base, or_span_t *spans,
                     uint64_t cap, uint64_t *position, int *err, uint64_t *err_offset)
{
    uint64_t pos = 0, n = 0;
    *err = 0;
    *err_offset = 0;
    while (pos < len && n < cap) {
        const uint8_t *rest = buf + pos;
        uint64_t rl = len - pos;
        if (rl < 12) break;
        uint64_t hl = 12 + (uint64_t)rd16(rest + 10);
        if (hl > rl) break;
        uint32_t bs, hs;
        int e = or_parse_header(rest, rl, &bs, &hs);
        if (e) { *err = e; *err_offset = base + pos; break; }
        if (bs > rl) break;
        uint32_t crc = rd32(rest + bs - 8);
        uint32_t isz = rd32(rest + bs - 4);
        if (isz > 65536) { *err = OR_BLK_ISIZE_TOO_BIG; *err_offset = base + pos; break; }
        spans[n].compressed_offset = base + pos;
        spans[n].payload_start = pos + hs;
        spans[n].payload_len = bs - hs - 8;
        spans[n].isize = isz;
        spans[n].crc32 = crc;
        ++n;
        pos += bs;
    }
    *position = pos;
    return n;
}

/*
 * A batch through the codec contract: plan checks (crates/fritillaria-cuda/src/backend.rs:95-133),
 * inflate, then verify in the order status -> size -> crc (backend.rs:221-256).
 * `out` must hold sum(isize) bytes; offsets gets n+1 entries.
 * Returns an or_batch_err_t kind (0 ok) and fills *e.
 */
int or_inflate_batch(const uint8_t *batch, uint64_t batch_len, const or_span_t *spans, uint64_t n,
                     uint8_t *out, uint64_t *offsets, or_error_t *e)
{
    memset(e, 0, sizeof *e);
    uint64_t total = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (spans[i].isize > 65536) {
            e->kind = OR_E_INVALID_BLOCK; e->offset = spans[i].compressed_offset; e->code = 1;
            return e->kind;
        }
        if (spans[i].payload_start + spans[i].payload_len > batch_len ||
            spans[i].payload_start + spans[i].payload_len < spans[i].payload_start) {
            e->kind = OR_E_INVALID_BLOCK; e->offset = spans[i].compressed_offset; e->code = 2;
            return e->kind;
        }
        total += spans[i].isize;
    }
    /* every member is inflated first, then verified in member order, as the
       device path does (the kernel runs all members, the host then walks them). */
    uint32_t *status = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *produced = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *crc = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t off = 0;
    for (uint64_t i = 0; i < n; ++i) {
        status[i] = (uint32_t)or_inflate_member(batch + spans[i].payload_start,
                                                (uint32_t)spans[i].payload_len, out + off,
                                                spans[i].isize, &produced[i]);
        crc[i] = status[i] == 0 ? or_crc32(out + off, produced[i]) : 0;
        off += spans[i].isize;
    }
    offsets[0] = 0;
    uint64_t cursor = 0;
    int kind = 0;
    for (uint64_t i = 0; i < n; ++i) {
        if (status[i] != 0) {
            kind = OR_E_INFLATE; e->offset = spans[i].compressed_offset; e->code = status[i];
            break;
        }
        if (produced[i] != spans[i].isize) {
            kind = OR_E_SIZE_MISMATCH; e->offset = spans[i].compressed_offset;
            e->expected = spans[i].isize; e->actual = produced[i];
            break;
        }
        if (crc[i] != spans[i].crc32) {
            kind = OR_E_CHECKSUM_MISMATCH; e->offset = spans[i].compressed_offset;
            e->expected = spans[i].crc32; e->actual = crc[i];
            break;
        }
        cursor += produced[i];
        offsets[i + 1] = cursor;
    }
    free(status);
    free(produced);
    free(crc);
    e->kind = kind;
    (void)total;
    return kind;
}

/* ------------------------------------------------------------------------ */
/* BAM.                                                                      */
/* ------------------------------------------------------------------------ */

/* crates/fritillaria-bam/src/columnar/header.rs:43-103. Returns 0 or an error
   code; *records_start on success, *err_pos on failure. */
int or_bam_parse_header(const uint8_t *buf, uint64_t len, uint64_t *records_start, uint64_t *err_pos,
                        uint32_t *n_ref)
{
    *err_pos = 0;
    if (len < 12) return 1;
    if (memcmp(buf, "BAM\1", 4) != 0) return 2;
    int32_t l_text = (int32_t)rd32(buf + 4);
    if (l_text < 0) { *err_pos = 4; return 3; }
    uint64_t text_end = 8 + (uint64_t)l_text;
    if (text_end > len) { *err_pos = 4; return 4; }
    if (text_end + 4 > len) { *err_pos = text_end; return 5; }
    int32_t nref = (int32_t)rd32(buf + text_end);
    if (nref < 0) { *err_pos = text_end; return 6; }
    uint64_t pos = text_end + 4;
    for (int32_t i = 0; i < nref; ++i) {
        if (pos + 4 > len) { *err_pos = pos; return 7; }
        int32_t l_name = (int32_t)rd32(buf + pos);
        if (l_name < 0) { *err_pos = pos; return 8; }
        pos += 4;
        if (pos + (uint64_t)l_name > len) { *err_pos = pos; return 9; }
        pos += (uint64_t)l_name;
        if (pos + 4 > len) { *err_pos = pos; return 10; }
        pos += 4;
    }
    *records_start = pos;
    *n_ref = (uint32_t)nref;
    return 0;
}

/* crates/fritillaria-bam/src/columnar/record.rs:58-95: the plain serial chain.
   Returns 0, or 1 (block_size < 32; *tail = position of the bad record). */
int or_bam_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets,
                        uint64_t cap, uint64_t *n, uint64_t *tail)
{
    uint64_t pos = start, k = 0;
    for (;;) {
        if (pos + 4 > len || pos + 4 < pos) break;
        uint32_t bs = rd32(buf + pos);
        if (bs < 32) { *n = k; *tail = pos; return 1; }
        uint64_t end = pos + 4 + bs;
        if (end > len) break;
        if (k < cap) offsets[k] = pos;
        ++k;
        pos = end;
    }
    *n = k;
    *tail = pos;
    return 0;
}

/* blocked.rs:106-133 (fields_fit). */
static int fields_fit(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t end)
{
    if (start + 24 > len) return 0;
    uint32_t name_len = buf[start + 12];
    if (name_len == 0) return 0;
    uint64_t n_cigar = rd16(buf + start + 16);
    uint64_t l_seq = rd32(buf + start + 20);
    uint64_t need = 36 + name_len + 4 * n_cigar + (l_seq + 1) / 2 + l_seq;
    return start + need <= end;
}

typedef struct {
    uint64_t count, land;
    int complete, valid;
} seg_t;

/* blocked.rs:141-187 (walk_segment). */
static seg_t walk_segment(const uint8_t *buf, uint64_t len, uint64_t from, uint64_t until)
{
    seg_t s = {0, from, 1, 1};
    uint64_t cur = from;
    for (;;) {
        if (cur >= until) { s.land = cur; s.complete = 1; return s; }
        if (cur + 4 > len) { s.land = cur; s.complete = 0; return s; }
        uint32_t bs = rd32(buf + cur);
        if (bs < 32) { s.land = cur; s.valid = 0; return s; }
        uint64_t end = cur + 4 + (uint64_t)bs;
        if (end > len) { s.land = cur; s.complete = 0; return s; }
        if (!fields_fit(buf, len, cur, end)) { s.land = cur; s.valid = 0; return s; }
        s.count++;
        cur = end;
    }
}

/*
 * blocked.rs:192-282 (scan_records_blocked), the semantics of the device scan:
 * per-block speculation, serial reconcile, emit. The result (offsets, tail,
 * error position) is what the device kernels must reproduce; the decomposition
 * itself is not a parity target.
 * Returns 0, or 1 = malformed with *tail = the reconcile's `cur`, or 2 = bad blocking.
 * *fallbacks counts blocks the chain did not enter at their start.
 */
int or_bam_scan_blocked(const uint8_t *buf, uint64_t len, const uint64_t *block_starts, uint64_t nb,
                        uint64_t start, uint64_t *offsets, uint64_t cap, uint64_t *n,
                        uint64_t *tail, uint64_t *fallbacks)
{
    *n = 0;
    *fallbacks = 0;
    if (block_starts[0] != 0 || block_starts[nb] != len) return 2;
    uint64_t cur = start, emitted = 0, t = len;
    int truncated = 0;
    for (uint64_t b = 0; b < nb; ++b) {
        if (truncated || cur >= block_starts[b + 1]) continue;
        if (cur != block_starts[b]) (*fallbacks)++;
        seg_t s = walk_segment(buf, len, cur, block_starts[b + 1]);
        if (!s.valid) { *tail = cur; *n = emitted; return 1; }
        uint64_t p = cur;
        for (uint64_t i = 0; i < s.count; ++i) {
            if (emitted + i < cap) offsets[emitted + i] = p;
            p += 4 + (uint64_t)rd32(buf + p);
        }
        emitted += s.count;
        cur = s.land;
        if (!s.complete) { t = s.land; truncated = 1; }
    }
    if (!truncated) t = cur;
    *n = emitted;
    *tail = t;
    return 0;
}

/*
 * crates/fritillaria-cuda/kernels/bam_decode.cu:254-291 and
 * crates/fritillaria-bam/src/columnar/record.rs:120-250: the 14 reference columns.
 * cols points at 14 arrays in the WIDTHS order of
 * crates/fritillaria-bam/src/columnar/device.rs:95-110.
 */
void or_bam_decode_columns(const uint8_t *buf, const uint64_t *offsets, uint64_t n,
                           int32_t *ref_id, int32_t *pos, uint16_t *flags, uint8_t *mapq,
                           uint32_t *l_seq, int32_t *next_ref_id, int32_t *next_pos, int32_t *tlen,
                           uint64_t *rec_off, uint32_t *cigar_start, uint32_t *seq_start,
                           uint32_t *qual_start, uint32_t *aux_start, uint32_t *rec_end)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t *r = buf + offsets[i];
        uint32_t bs = rd32(r);
        uint32_t name_len = r[12];
        uint32_t nc = rd16(r + 16);
        uint32_t ls = rd32(r + 20);
        ref_id[i] = (int32_t)rd32(r + 4);
        pos[i] = (int32_t)rd32(r + 8);
        mapq[i] = r[13];
        flags[i] = rd16(r + 18);
        l_seq[i] = ls;
        next_ref_id[i] = (int32_t)rd32(r + 24);
        next_pos[i] = (int32_t)rd32(r + 28);
        tlen[i] = (int32_t)rd32(r + 32);
        rec_off[i] = offsets[i];
        uint32_t c = 36 + name_len;
        uint32_t s = c + 4 * nc;
        uint32_t q = s + (ls + 1) / 2;
        uint32_t a = q + ls;
        cigar_start[i] = c;
        seq_start[i] = s;
        qual_start[i] = q;
        aux_start[i] = a;
        rec_end[i] = 4 + bs;
    }
}

/*
 * Virtual offset of an uncompressed batch position: (coffset << 16) | uoffset of
 * the member holding it, where a position on a member boundary belongs to the
 * next non-empty member. crates/fritillaria-core/src/virtual_offset.rs:24-48,
 * crates/fritillaria-bgzf/src/read.rs:118-148, io/block.rs:34-46.
 * block_offsets has nb+1 entries (prefix sum of ISIZE); coffsets has nb.
 */
uint64_t or_virtual_offset(const uint64_t *block_offsets, const uint64_t *coffsets, uint64_t nb,
                           uint64_t p)
{
    /* partition_point(|o| o <= p) - 1 */
    uint64_t lo = 0, hi = nb + 1;
    while (lo < hi) {
        uint64_t mid = (lo + hi) / 2;
        if (block_offsets[mid] <= p) lo = mid + 1; else hi = mid;
    }
    uint64_t b = lo - 1;
    if (b >= nb) b = nb - 1; /* p == total: end of the last member */
    return (coffsets[b] << 16) | (p - block_offsets[b]);
}

/* ------------------------------------------------------------------------ */
/* FASTQ record boundaries and line bounds.                                  */
/* ------------------------------------------------------------------------ */

/* Index of the next '\n' at or after pos, or len when there is none. */
static uint64_t fq_line_end(const uint8_t *buf, uint64_t len, uint64_t pos)
{
    if (pos >= len) return len;
    const uint8_t *p = (const uint8_t *)memchr(buf + pos, '\n', (size_t)(len - pos));
    return p ? (uint64_t)(p - buf) : len;
}

/*
 * The four line bounds of a record at pos, relative to pos, or 0 when pos does
 * not begin a whole record. crates/fritillaria-fastq/src/columnar/record.rs:85-131:
 * '@' at pos; newlines end the definition, sequence and '+' lines; '+' opens
 * line 3; the quality line runs to the next newline or the buffer end and must
 * be as long as the sequence line; the record ends past the quality newline
 * when there is one.
 */
int or_fastq_bounds(const uint8_t *buf, uint64_t len, uint64_t pos, uint32_t b[5])
{
    if (pos >= len || buf[pos] != '@') return 0;
    uint64_t def_end = fq_line_end(buf, len, pos);
    if (def_end >= len) return 0;
    uint64_t seq_start = def_end + 1;
    uint64_t seq_end = fq_line_end(buf, len, seq_start);
    if (seq_end >= len) return 0;
    uint64_t plus_start = seq_end + 1;
    if (plus_start >= len || buf[plus_start] != '+') return 0;
    uint64_t plus_end = fq_line_end(buf, len, plus_start);
    if (plus_end >= len) return 0;
    uint64_t qual_start = plus_end + 1;
    uint64_t seq_len = seq_end - seq_start;
    uint64_t qual_end = fq_line_end(buf, len, qual_start);
    if (qual_end - qual_start != seq_len) return 0;
    uint64_t end = qual_end < len ? qual_end + 1 : qual_end;
    b[0] = (uint32_t)(seq_start - pos);
    b[1] = (uint32_t)(plus_start - pos);
    b[2] = (uint32_t)(qual_start - pos);
    b[3] = (uint32_t)seq_len;
    b[4] = (uint32_t)(end - pos);
    return 1;
}

/*
 * Serial record walk, crates/fritillaria-fastq/src/columnar/record.rs:145-176:
 * newlines between records are skipped; a position that is not a whole record
 * ends the walk (the tail) when it holds '@', and is a Malformed error
 * (err_pos = that position) otherwise. Returns 0 or OR_E_MALFORMED.
 * bounds (5 per record) may be NULL.
 */
int or_fastq_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets,
                          uint32_t *bounds, uint64_t cap, uint64_t *n, uint64_t *tail, uint64_t *err_pos)
{
    uint64_t pos = start, k = 0;
    *n = 0;
    while (pos < len) {
        if (buf[pos] == '\n') {
            ++pos;
            continue;
        }
        uint32_t b[5];
        if (!or_fastq_bounds(buf, len, pos, b)) {
            if (buf[pos] == '@') break;
            *err_pos = pos;
            *tail = pos;
            return OR_E_MALFORMED;
        }
        if (k < cap) {
            offsets[k] = pos;
            if (bounds) memcpy(bounds + 5 * k, b, sizeof b);
        }
        ++k;
        pos += b[4];
    }
    *n = k;
    *tail = pos;
    return 0;
}

/* Number of line starts at or after start (start itself and every byte after
   a '\n' below len) that begin a whole record: the survivors of the sieve of
   crates/fritillaria-fastq/src/columnar/speculative.rs:96-126. */
uint64_t or_fastq_survivors(const uint8_t *buf, uint64_t len, uint64_t start)
{
    uint64_t s = 0;
    uint32_t b[5];
    for (uint64_t i = start; i < len; ++i)
        if ((i == start || buf[i - 1] == '\n') && or_fastq_bounds(buf, len, i, b)) ++s;
    return s;
}

/* ------------------------------------------------------------------------ */
/* BCF record boundaries, typed values and site-core columns.               */
/* ------------------------------------------------------------------------ */

static uint32_t bcf_type_size(uint32_t code)
{
    switch (code) {
    case 0: return 0;
    case 1: case 7: return 1;
    case 2: return 2;
    case 3: case 5: return 4;
    default: return 0xFFFFFFFFu;
    }
}

/* Header bytes and element count of the typed value at pos of buf[0..len), or 0
   when it is not one (crates/fritillaria-bcf/src/columnar/typed.rs:255-300):
   reserved type codes fail; an escaped count (nibble 15) is an atomic integer
   value of at least 15 that follows the descriptor. */
static uint32_t bcf_descriptor(const uint8_t *buf, uint64_t len, uint64_t pos, uint64_t *count, uint32_t *kind)
{
    if (pos >= len) return 0;
    uint32_t d = buf[pos], code = d & 15;
    if (bcf_type_size(code) == 0xFFFFFFFFu) return 0;
    *kind = code;
    if ((d >> 4) < 15) {
        *count = d >> 4;
        return 1;
    }
    if (pos + 1 >= len) return 0;
    uint32_t in = buf[pos + 1], ic = in & 15;
    if ((in >> 4) != 1 || (ic != 1 && ic != 2 && ic != 3)) return 0;
    uint32_t w = bcf_type_size(ic);
    if (pos + 2 + w > len) return 0;
    const uint8_t *p = buf + pos + 2;
    int32_t v = ic == 1 ? (int32_t)(int8_t)p[0] : ic == 2 ? (int32_t)(int16_t)rd16(p) : (int32_t)rd32(p);
    if (v < 15) return 0;
    *count = (uint64_t)v;
    return 2 + w;
}

/* End of the typed value at pos within buf[0..len), or 0 (typed.rs skip). */
static uint64_t bcf_skip(const uint8_t *buf, uint64_t len, uint64_t pos)
{
    uint64_t count = 0;
    uint32_t kind = 0;
    uint32_t h = bcf_descriptor(buf, len, pos, &count, &kind);
    if (!h) return 0;
    uint64_t end = pos + h + count * bcf_type_size(kind);
    return end > len ? 0 : end;
}

/* looks_like_a_record (crates/fritillaria-bcf/src/columnar/record.rs:96-160). */
int or_bcf_looks_like(const uint8_t *buf, uint64_t len, uint64_t pos, uint32_t samples, uint32_t contigs)
{
    if (pos + 8 > len) return 0;
    uint32_t ls = rd32(buf + pos), li = rd32(buf + pos + 4);
    if (ls < 24) return 0;
    uint64_t end = pos + 8 + (uint64_t)ls + li;
    if (end > len || pos + 32 > len) return 0;
    int32_t chrom = (int32_t)rd32(buf + pos + 8), p = (int32_t)rd32(buf + pos + 12), rlen = (int32_t)rd32(buf + pos + 16);
    uint32_t n_allele = rd16(buf + pos + 26);
    uint32_t n_sample = buf[pos + 28] | (buf[pos + 29] << 8) | (buf[pos + 30] << 16);
    uint32_t n_fmt = buf[pos + 31];
    if (chrom < 0 || (uint32_t)chrom >= contigs) return 0;
    if (p < -1 || rlen < 0) return 0;
    if (n_sample != samples || n_allele == 0) return 0;
    if (n_fmt == 0) return li == 0;
    return (uint64_t)li >= 2ull * n_fmt;
}

/* Serial record walk (record.rs scan_records): fewer than 8 bytes left or a
   record past the end is the tail; l_shared < 24 is Malformed at that record. */
int or_bcf_scan_records(const uint8_t *buf, uint64_t len, uint64_t start, uint64_t *offsets, uint64_t cap,
                        uint64_t *n, uint64_t *tail, uint64_t *err_pos)
{
    uint64_t pos = start, k = 0;
    for (;;) {
        if (pos + 8 > len) break;
        uint32_t ls = rd32(buf + pos), li = rd32(buf + pos + 4);
        if (ls < 24) {
            *n = k;
            *tail = pos;
            *err_pos = pos;
            return OR_E_MALFORMED;
        }
        uint64_t end = pos + 8 + (uint64_t)ls + li;
        if (end > len) break;
        if (k < cap) offsets[k] = pos;
        ++k;
        pos = end;
    }
    *n = k;
    *tail = pos;
    return 0;
}

/* RecordBatch::decode columns (crates/fritillaria-bcf/src/columnar/batch.rs:158-172,
   record.rs Record::new / bounds): fixed fields of the site core, then ID,
   n_allele alleles and FILTER skipped as typed values inside the record. Returns
   0, or OR_E_MALFORMED with *bad = index of the first record that fails. */
int or_bcf_decode_columns(const uint8_t *buf, const uint64_t *offsets, uint64_t n, uint64_t tail,
                          int32_t *chrom, int32_t *pos, int32_t *rlen, uint32_t *qual, uint16_t *n_info,
                          uint16_t *n_allele, uint32_t *n_sample, uint8_t *n_fmt, uint64_t *rec_off,
                          uint32_t *alleles_start, uint32_t *filters_start, uint32_t *info_start,
                          uint32_t *genotypes_start, uint32_t *rec_end, uint64_t *bad)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t o = offsets[i], e = i + 1 < n ? offsets[i + 1] : tail;
        const uint8_t *r = buf + o;
        const uint64_t L = e - o;
        uint32_t ls = rd32(r), li = rd32(r + 4);
        if (L < 32 || ls < 24 || 8ull + ls + li != L) {
            *bad = i;
            return OR_E_MALFORMED;
        }
        chrom[i] = (int32_t)rd32(r + 8);
        pos[i] = (int32_t)rd32(r + 12);
        rlen[i] = (int32_t)rd32(r + 16);
        qual[i] = rd32(r + 20);
        n_info[i] = rd16(r + 24);
        n_allele[i] = rd16(r + 26);
        n_sample[i] = r[28] | (r[29] << 8) | (r[30] << 16);
        n_fmt[i] = r[31];
        rec_off[i] = o;
        uint64_t a = bcf_skip(r, L, 32);
        if (!a) { *bad = i; return OR_E_MALFORMED; }
        uint64_t f = a;
        for (uint32_t k = 0; k < n_allele[i]; ++k) {
            f = bcf_skip(r, L, f);
            if (!f) { *bad = i; return OR_E_MALFORMED; }
        }
        uint64_t inf = bcf_skip(r, L, f);
        if (!inf) { *bad = i; return OR_E_MALFORMED; }
        alleles_start[i] = (uint32_t)a;
        filters_start[i] = (uint32_t)f;
        info_start[i] = (uint32_t)inf;
        genotypes_start[i] = 8 + ls;
        rec_end[i] = (uint32_t)L;
    }
    return 0;
}

/* Record::validate (record.rs:292-330): every site field and every FORMAT field
   walks exactly to its block's end. 1 = valid. */
int or_bcf_validate(const uint8_t *buf, uint64_t len, uint64_t pos)
{
    if (pos + 32 > len) return 0;
    uint64_t ls = rd32(buf + pos), li = rd32(buf + pos + 4);
    uint64_t indiv = pos + 8 + ls, end = indiv + li;
    if (ls < 24 || end > len) return 0;
    uint32_t n_info = rd16(buf + pos + 24), n_allele = rd16(buf + pos + 26);
    uint32_t n_sample = buf[pos + 28] | (buf[pos + 29] << 8) | (buf[pos + 30] << 16), n_fmt = buf[pos + 31];
    uint64_t at = bcf_skip(buf, indiv, pos + 32);
    if (!at) return 0;
    for (uint32_t i = 0; i < n_allele; ++i) if (!(at = bcf_skip(buf, indiv, at))) return 0;
    if (!(at = bcf_skip(buf, indiv, at))) return 0;
    for (uint32_t i = 0; i < n_info; ++i) {
        if (!(at = bcf_skip(buf, indiv, at))) return 0;
        if (!(at = bcf_skip(buf, indiv, at))) return 0;
    }
    if (at != indiv) return 0;
    for (uint32_t i = 0; i < n_fmt; ++i) {
        if (!(at = bcf_skip(buf, end, at))) return 0;
        uint64_t count = 0;
        uint32_t kind = 0;
        uint32_t h = bcf_descriptor(buf, end, at, &count, &kind);
        if (!h) return 0;
        at += h + count * bcf_type_size(kind) * n_sample;
        if (at > end) return 0;
    }
    return at == end;
}

/* ------------------------------------------------------------------------ */
/* BAM structure-of-arrays extras: resolved CIGAR ops, packed sequence, aux  */
/* tag index.                                                                */
/* ------------------------------------------------------------------------ */

/* One aux field at p of the record's aux block [.., end) (crates/fritillaria-bam/src/columnar/aux.rs
   read_value / read_array / Fields): writes the entry and returns the next
   field's offset, or 0 when the field is malformed. Entry: tag[2], type,
   subtype ('B' only), value offset (relative to the record), element count
   (1 for scalars, the string length without its NUL for Z/H, the array count
   for B), value bytes (NUL included for Z/H). */
static uint64_t aux_field(const uint8_t *r, uint64_t p, uint64_t end, or_aux_entry_t *e)
{
    if (p + 3 > end) return 0;
    const uint8_t ty = r[p + 2];
    uint64_t v = p + 3, size, count = 1;
    uint8_t sub = 0;
    switch (ty) {
    case 'A': case 'c': case 'C': size = 1; break;
    case 's': case 'S': size = 2; break;
    case 'i': case 'I': case 'f': size = 4; break;
    case 'Z': case 'H': {
        uint64_t q = v;
        while (q < end && r[q] != 0) ++q;
        if (q >= end) return 0;
        count = q - v;
        size = count + 1;
        if (ty == 'H') {
            if (count % 2) return 0;
            for (uint64_t k = v; k < q; ++k) {
                uint8_t c = r[k];
                if (!((c >= '0' && c <= '9') || (c >= 'a' && c <= 'f') || (c >= 'A' && c <= 'F'))) return 0;
            }
        }
        break;
    }
    case 'B': {
        if (v + 5 > end) return 0;
        sub = r[v];
        uint32_t w = sub == 'c' || sub == 'C' ? 1 : sub == 's' || sub == 'S' ? 2
                   : sub == 'i' || sub == 'I' || sub == 'f' ? 4 : 0;
        if (!w) return 0;
        count = rd32(r + v + 1);
        v += 5;
        size = count * w;
        break;
    }
    default: return 0;
    }
    if (v + size > end) return 0;
    if (e) {
        e->tag[0] = r[p];
        e->tag[1] = r[p + 1];
        e->type = ty;
        e->subtype = sub;
        e->value_offset = (uint32_t)v;
        e->count = (uint32_t)count;
        e->size = (uint32_t)size;
    }
    return v + size;
}

/*
 * Per record (offsets from a scan): the resolved CIGAR (Record::cigar_resolved,
 * crates/fritillaria-bam/src/columnar/record.rs:316-361: the placeholder
 * "<l_seq>S <n>N" with exactly two ops is replaced by the CG:B:I array, which
 * must exist), the packed 4-bit sequence and the aux fields. Counts go to
 * n_ops / seq_bytes / n_aux (per record); with non-NULL outputs the data are
 * written at the given prefix offsets. Returns 0, or OR_E_MALFORMED with
 * *bad = first failing record.
 */
int or_bam_soa(const uint8_t *buf, const uint64_t *offsets, uint64_t n, uint32_t *n_ops, uint32_t *seq_bytes,
               uint32_t *n_aux, const uint64_t *op_off, uint32_t *ops, const uint64_t *seq_off, uint8_t *seq,
               const uint64_t *aux_off, or_aux_entry_t *aux, uint64_t *bad)
{
    for (uint64_t i = 0; i < n; ++i) {
        const uint8_t *r = buf + offsets[i];
        const uint64_t end = 4 + (uint64_t)rd32(r);
        const uint32_t name_len = r[12], nc = rd16(r + 16), ls = rd32(r + 20);
        const uint64_t c0 = 36 + name_len, s0 = c0 + 4ull * nc, q0 = s0 + (ls + 1) / 2, a0 = q0 + ls;
        uint32_t k = 0;
        uint64_t cg = 0, cg_count = 0;
        int cg_ok = 0;
        for (uint64_t p = a0; p < end;) {
            or_aux_entry_t e;
            uint64_t nx = aux_field(r, p, end, &e);
            if (!nx) { *bad = i; return OR_E_MALFORMED; }
            if (aux) aux[aux_off[i] + k] = e;
            if (e.tag[0] == 'C' && e.tag[1] == 'G' && !cg) {
                cg = e.value_offset;
                cg_count = e.count;
                cg_ok = e.type == 'B' && e.subtype == 'I';
                cg = cg ? cg : 1;
            }
            ++k;
            p = nx;
        }
        n_aux[i] = k;
        int placeholder = 0;
        if (nc == 2) {
            uint32_t o0 = rd32(r + c0), o1 = rd32(r + c0 + 4);
            placeholder = (o0 & 15) == 4 && (o0 >> 4) == ls && (o1 & 15) == 3;
        }
        if (placeholder) {
            if (!cg || !cg_ok) { *bad = i; return OR_E_MALFORMED; }
            n_ops[i] = (uint32_t)cg_count;
            if (ops) for (uint64_t j = 0; j < cg_count; ++j) ops[op_off[i] + j] = rd32(r + cg + 4 * j);
        } else {
            n_ops[i] = nc;
            if (ops) for (uint32_t j = 0; j < nc; ++j) ops[op_off[i] + j] = rd32(r + c0 + 4ull * j);
        }
        seq_bytes[i] = (ls + 1) / 2;
        if (seq) memcpy(seq + seq_off[i], r + s0, (ls + 1) / 2);
    }
    return 0;
}
