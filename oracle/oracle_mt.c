/*
 * oracle_mt.c -- host-core CPU baseline (test/bench infrastructure only).
 *
 * Restates the structure of noodles' multithreaded BGZF reader feeding the BAM
 * record reader: one framing walk over member headers
 * (crates/fritillaria-bgzf/src/io/multithreaded_reader.rs:141-153,
 *  io/reader/frame.rs:10-37), a pool of workers each inflating whole members and
 * checking CRC32 + ISIZE (frame.rs:103-150), an ordered hand-off
 * (multithreaded_reader.rs:359-393) to one thread that walks `block_size`
 * records (crates/fritillaria-bam/src/io/reader/record.rs:7-22) and fills the
 * column set of crates/fritillaria-bam/src/columnar/batch.rs:165-180.
 * Inflate and CRC use system zlib 1.2.11 in place of zlib-rs 0.6.7.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

typedef struct {
    const uint8_t *file;
    or_span_t *spans;
    uint64_t n;
    uint64_t *out_off;
    uint8_t *out;
    atomic_uint_fast64_t next;
    atomic_int *done; /* 1 ok, 2 failed */
} mt_ctx_t;

static int inflate_raw(const uint8_t *in, uint32_t in_len, uint8_t *out, uint32_t cap, uint32_t *produced)
{
    z_stream zs;
    memset(&zs, 0, sizeof zs);
    if (inflateInit2(&zs, -15) != Z_OK) return -1;
    zs.next_in = (Bytef *)in;
    zs.avail_in = in_len;
    zs.next_out = out;
    zs.avail_out = cap;
    int r = inflate(&zs, Z_FINISH);
    *produced = (uint32_t)zs.total_out;
    inflateEnd(&zs);
    return r == Z_STREAM_END ? 0 : -1;
}

static void *worker(void *arg)
{
    mt_ctx_t *c = (mt_ctx_t *)arg;
    for (;;) {
        uint64_t i = atomic_fetch_add(&c->next, 1);
        if (i >= c->n) break;
        const or_span_t *s = &c->spans[i];
        uint32_t produced = 0;
        int ok = inflate_raw(c->file + s->payload_start, (uint32_t)s->payload_len,
                             c->out + c->out_off[i], s->isize, &produced) == 0 &&
                 produced == s->isize &&
                 (uint32_t)crc32(0L, c->out + c->out_off[i], produced) == s->crc32;
        atomic_store_explicit(&c->done[i], ok ? 1 : 2, memory_order_release);
    }
    return NULL;
}

static uint32_t rd32(const uint8_t *p)
{
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24);
}

/* Waits until every member covering [0, upto) has been inflated. */
static int wait_upto(mt_ctx_t *c, uint64_t *ready_members, uint64_t upto)
{
    while (*ready_members < c->n && c->out_off[*ready_members] < upto) {
        int d;
        while ((d = atomic_load_explicit(&c->done[*ready_members], memory_order_acquire)) == 0) {
        }
        if (d != 1) return -1;
        (*ready_members)++;
    }
    return 0;
}

/*
 * Decodes a whole BGZF BAM held in memory with `nthreads` inflate workers.
 * Fills res[0]=records, res[1]=inflated bytes, res[2]=sum(pos), res[3]=sum(flag),
 * res[4]=sum(l_seq), res[5]=sum(record offsets), res[6]=members.
 * With soa != 0 every record's resolved CIGAR ops, packed sequence and aux
 * entries are also copied into growing columns (or_bam_soa per record, the
 * accessors of crates/fritillaria-bam/src/columnar/record.rs:220-361), and
 * res[7]=total ops, res[8]=sum of op words, res[9]=packed sequence bytes,
 * res[10]=sum of those bytes, res[11]=aux entries, res[12]=sum of their value
 * offsets. res must hold 13 words.
 * Returns 0, or <0 on a framing / inflate / record / aux error.
 */
/* Grows *p (elements of `size` bytes) to hold at least `need` elements. */
static int grow(void **p, uint64_t *cap, uint64_t need, size_t size)
{
    if (need <= *cap) return 0;
    uint64_t c = *cap ? *cap : 1024;
    while (c < need) c *= 2;
    void *q = realloc(*p, c * size);
    if (!q) return -1;
    *p = q;
    *cap = c;
    return 0;
}

int or_mt_bam_columns(const uint8_t *file, uint64_t len, int nthreads, int soa, uint64_t *res)
{
    uint64_t cap = len / 28 + 2;
    or_span_t *spans = (or_span_t *)malloc(sizeof(or_span_t) * cap);
    uint64_t position = 0, eoff = 0;
    int err = 0;
    uint64_t n = or_discover(file, len, 0, spans, cap, &position, &err, &eoff);
    if (err || position != len) { free(spans); return -1; }
    uint64_t *out_off = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    out_off[0] = 0;
    for (uint64_t i = 0; i < n; ++i) out_off[i + 1] = out_off[i] + spans[i].isize;
    uint64_t total = out_off[n];
    uint8_t *out = (uint8_t *)malloc(total ? total : 1);
    atomic_int *done = (atomic_int *)calloc(n ? n : 1, sizeof(atomic_int));
    mt_ctx_t c = {file, spans, n, out_off, out, 0, done};
    atomic_init(&c.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &c);

    /* Ordered consumer: header, then the record walk with column fill. */
    int rc = 0;
    uint64_t ready = 0;
    uint64_t recs = 0, sum_pos = 0, sum_flag = 0, sum_lseq = 0, sum_off = 0;
    uint64_t col_cap = total / 36 + 1;
    int32_t *c_ref = (int32_t *)malloc(4 * col_cap), *c_pos = (int32_t *)malloc(4 * col_cap);
    uint16_t *c_flag = (uint16_t *)malloc(2 * col_cap);
    uint8_t *c_mapq = (uint8_t *)malloc(col_cap);
    uint32_t *c_lseq = (uint32_t *)malloc(4 * col_cap);
    int32_t *c_nref = (int32_t *)malloc(4 * col_cap), *c_npos = (int32_t *)malloc(4 * col_cap);
    int32_t *c_tlen = (int32_t *)malloc(4 * col_cap);
    uint64_t *c_off = (uint64_t *)malloc(8 * col_cap);
    /* SoA extras: columns grown as records arrive */
    uint32_t *s_ops = NULL;
    uint8_t *s_seq = NULL;
    or_aux_entry_t *s_aux = NULL;
    uint64_t cap_ops = 0, cap_seq = 0, cap_aux = 0, n_ops_tot = 0, n_seq_tot = 0, n_aux_tot = 0;
    uint64_t sum_ops = 0, sum_seq = 0, sum_auxoff = 0;
    if (wait_upto(&c, &ready, 12) != 0) { rc = -2; goto join; }
    {
        uint64_t need = 12;
        if (total < 12 || memcmp(out, "BAM\1", 4) != 0) { rc = -3; goto join; }
        uint64_t l_text = rd32(out + 4);
        need = 8 + l_text + 4;
        if (wait_upto(&c, &ready, need) != 0 || need > total) { rc = -3; goto join; }
        uint32_t nref = rd32(out + 8 + l_text);
        uint64_t pos = 8 + l_text + 4;
        for (uint32_t i = 0; i < nref; ++i) {
            if (wait_upto(&c, &ready, pos + 4) != 0 || pos + 4 > total) { rc = -3; goto join; }
            uint64_t l_name = rd32(out + pos);
            pos += 4 + l_name + 4;
        }
        for (;;) {
            if (wait_upto(&c, &ready, pos + 4) != 0) { rc = -2; goto join; }
            if (pos + 4 > total) break;
            uint32_t bs = rd32(out + pos);
            if (bs < 32) { rc = -4; goto join; }
            uint64_t end = pos + 4 + bs;
            if (wait_upto(&c, &ready, end) != 0) { rc = -2; goto join; }
            if (end > total) { rc = -4; goto join; }
            const uint8_t *r = out + pos;
            c_ref[recs] = (int32_t)rd32(r + 4);
            c_pos[recs] = (int32_t)rd32(r + 8);
            c_mapq[recs] = r[13];
            c_flag[recs] = (uint16_t)(r[18] | (r[19] << 8));
            c_lseq[recs] = rd32(r + 20);
            c_nref[recs] = (int32_t)rd32(r + 24);
            c_npos[recs] = (int32_t)rd32(r + 28);
            c_tlen[recs] = (int32_t)rd32(r + 32);
            c_off[recs] = pos;
            sum_pos += (uint64_t)(int64_t)c_pos[recs];
            sum_flag += c_flag[recs];
            sum_lseq += c_lseq[recs];
            sum_off += pos;
            if (soa) {
                /* a record of bs bytes holds at most bs/4 ops or aux fields and bs sequence bytes */
                if (grow((void **)&s_ops, &cap_ops, n_ops_tot + bs / 4 + 1, 4) ||
                    grow((void **)&s_seq, &cap_seq, n_seq_tot + bs + 1, 1) ||
                    grow((void **)&s_aux, &cap_aux, n_aux_tot + bs / 4 + 1, sizeof(or_aux_entry_t))) {
                    rc = -5;
                    goto join;
                }
                uint32_t k_ops, k_seq, k_aux;
                uint64_t bad = 0, at = pos;
                if (or_bam_soa(out, &at, 1, &k_ops, &k_seq, &k_aux, &n_ops_tot, s_ops, &n_seq_tot, s_seq,
                               &n_aux_tot, s_aux, &bad) != 0) {
                    rc = -6;
                    goto join;
                }
                for (uint32_t j = 0; j < k_ops; ++j) sum_ops += s_ops[n_ops_tot + j];
                for (uint32_t j = 0; j < k_seq; ++j) sum_seq += s_seq[n_seq_tot + j];
                for (uint32_t j = 0; j < k_aux; ++j) sum_auxoff += s_aux[n_aux_tot + j].value_offset;
                n_ops_tot += k_ops;
                n_seq_tot += k_seq;
                n_aux_tot += k_aux;
            }
            ++recs;
            pos = end;
        }
    }
join:
    if (rc != 0) {
        /* let the pool drain: workers only stop when the queue is empty */
    }
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    res[0] = recs; res[1] = total; res[2] = sum_pos; res[3] = sum_flag;
    res[4] = sum_lseq; res[5] = sum_off; res[6] = n;
    res[7] = n_ops_tot; res[8] = sum_ops; res[9] = n_seq_tot; res[10] = sum_seq;
    res[11] = n_aux_tot; res[12] = sum_auxoff;
    free(s_ops); free(s_seq); free(s_aux);
    free(th); free(done); free(out); free(out_off); free(spans);
    free(c_ref); free(c_pos); free(c_flag); free(c_mapq); free(c_lseq);
    free(c_nref); free(c_npos); free(c_tlen); free(c_off);
    return rc;
}

/*
 * Inflates a whole BGZF stream held in memory with `nthreads` workers (zlib,
 * CRC32 and ISIZE checked per member), the same member-parallel inflate the
 * BAM reader above uses, for formats whose record walk runs afterwards on the
 * whole buffer (the BCF and FASTQ baselines). *out_len receives the inflated
 * size; out must hold it (call with out == NULL to learn it). Returns 0, or <0
 * on a framing / inflate error.
 */
int or_mt_inflate(const uint8_t *file, uint64_t len, int nthreads, uint8_t *out, uint64_t *out_len)
{
    uint64_t cap = len / 28 + 2;
    or_span_t *spans = (or_span_t *)malloc(sizeof(or_span_t) * cap);
    uint64_t position = 0, eoff = 0;
    int err = 0;
    uint64_t n = or_discover(file, len, 0, spans, cap, &position, &err, &eoff);
    if (err || position != len) { free(spans); return -1; }
    uint64_t *out_off = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    out_off[0] = 0;
    for (uint64_t i = 0; i < n; ++i) out_off[i + 1] = out_off[i] + spans[i].isize;
    *out_len = out_off[n];
    if (!out) { free(out_off); free(spans); return 0; }
    atomic_int *done = (atomic_int *)calloc(n ? n : 1, sizeof(atomic_int));
    mt_ctx_t c = {file, spans, n, out_off, out, 0, done};
    atomic_init(&c.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t *th = (pthread_t *)malloc(sizeof(pthread_t) * nthreads);
    for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, worker, &c);
    for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    int rc = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (atomic_load(&done[i]) != 1) rc = -2;
    free(th); free(done); free(out_off); free(spans);
    return rc;
}
