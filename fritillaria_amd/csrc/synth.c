/*
 * synth.c -- synthetic BGZF inputs for tests and benchmarks.
 *
 * Writes BAM streams the way the vendored noodles writer frames them: the whole
 * uncompressed stream (header, then records back to back) is cut every 65,495
 * bytes with no regard for record boundaries
 * (crates/fritillaria-bgzf/src/io/writer.rs:17,24-25,246-255;
 *  crates/fritillaria-bam/src/io/writer.rs:123-126), each piece becomes one BGZF
 * member with an XLEN=6 `BC` header, and the file ends with the 28-byte EOF
 * member (crates/fritillaria-bgzf/src/block.rs:40-43).
 *
 * Every record is generated from a counter-based RNG keyed by (seed, index), so
 * records can be produced in parallel and the output is identical for any
 * thread count. Members are compressed in parallel with zlib raw deflate.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "synth.h"

static inline uint64_t splitmix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

typedef struct { uint64_t s; } rng_t;
static inline uint64_t rnext(rng_t *r) { r->s += 0x9E3779B97F4A7C15ull; return splitmix(r->s); }
static inline double runif(rng_t *r) { return (rnext(r) >> 11) * (1.0 / 9007199254740992.0); }
static inline uint32_t rbelow(rng_t *r, uint32_t n) { return (uint32_t)(((rnext(r) >> 32) * (uint64_t)n) >> 32); }
static inline double rnorm(rng_t *r)
{
    double u1 = runif(r), u2 = runif(r);
    if (u1 < 1e-300) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* Phred ~ round(N(35,5)) clipped to [2,41], sampled by inverse CDF over a
   4096-entry table (12 random bits per quality value). */
static uint8_t QTAB[4096];
static pthread_once_t qonce = PTHREAD_ONCE_INIT;
static void build_qtab(void)
{
    for (int i = 0; i < 4096; ++i) {
        double p = (i + 0.5) / 4096.0;
        /* inverse normal CDF by bisection on erfc */
        double lo = -10, hi = 10;
        for (int k = 0; k < 80; ++k) {
            double mid = 0.5 * (lo + hi);
            double c = 0.5 * erfc(-mid / sqrt(2.0));
            if (c < p) lo = mid; else hi = mid;
        }
        double q = 35.0 + 5.0 * 0.5 * (lo + hi);
        long v = lround(q);
        if (v < 2) v = 2;
        if (v > 41) v = 41;
        QTAB[i] = (uint8_t)v;
    }
}

/* SAM spec reg2bin (0-based, end exclusive). */
static int reg2bin(int64_t beg, int64_t end)
{
    --end;
    if (beg >> 14 == end >> 14) return (int)(((1 << 15) - 1) / 7 + (beg >> 14));
    if (beg >> 17 == end >> 17) return (int)(((1 << 12) - 1) / 7 + (beg >> 17));
    if (beg >> 20 == end >> 20) return (int)(((1 << 9) - 1) / 7 + (beg >> 20));
    if (beg >> 23 == end >> 23) return (int)(((1 << 6) - 1) / 7 + (beg >> 23));
    if (beg >> 26 == end >> 26) return (int)(((1 << 3) - 1) / 7 + (beg >> 26));
    return 0;
}

static inline void w32(uint8_t *p, uint32_t v) { p[0] = v; p[1] = v >> 8; p[2] = v >> 16; p[3] = v >> 24; }
static inline void w16(uint8_t *p, uint16_t v) { p[0] = v; p[1] = v >> 8; }

/* Shape of one record: everything that decides its size. */
typedef struct {
    uint32_t l_seq;
    uint32_t n_cigar;
    uint32_t cigar[3];
    uint32_t ref_span;
    uint16_t flag;
    uint8_t nm;
    uint32_t n_mm;      /* modified-base calls (long-read aux) */
    uint32_t mm_bytes;  /* bytes of the MM:Z value incl. NUL */
    uint64_t mm_seed;
} shape_t;

static const double GEOM_P = 0.05;

static void make_shape(const synth_bam_params *P, uint64_t i, shape_t *s)
{
    rng_t r = {splitmix(P->seed * 0x100000001B3ull ^ (i * 0xD6E8FEB86659FD93ull))};
    memset(s, 0, sizeof *s);
    uint32_t L;
    if (P->read_mode == 0) {
        L = P->len_fixed;
    } else if (P->read_mode == 1) {
        L = P->len_min + rbelow(&r, P->len_max - P->len_min + 1);
    } else {
        double v = exp(log(P->ln_median) + P->ln_sigma * rnorm(&r));
        if (v > P->len_cap || (P->cap_every && i % P->cap_every == P->cap_every / 2)) v = P->len_cap;
        if (v < 1) v = 1;
        L = (uint32_t)v;
    }
    s->l_seq = L;
    if (P->unaligned) {
        s->flag = 4;
        s->n_cigar = 0;
        s->ref_span = 0;
    } else {
        s->flag = (rnext(&r) >> 63) ? 16 : 0;
        if (P->indel_p > 0 && runif(&r) < P->indel_p && L >= 10) {
            uint32_t b = 1 + rbelow(&r, 3);
            int ins = (int)(rnext(&r) >> 63);
            uint32_t a = 1 + rbelow(&r, L - 5);
            if (ins) {
                uint32_t c = L - a - b;
                s->n_cigar = 3;
                s->cigar[0] = (a << 4) | 0;
                s->cigar[1] = (b << 4) | 1;
                s->cigar[2] = (c << 4) | 0;
                s->ref_span = a + c;
            } else {
                uint32_t c = L - a;
                s->n_cigar = 3;
                s->cigar[0] = (a << 4) | 0;
                s->cigar[1] = (b << 4) | 2;
                s->cigar[2] = (c << 4) | 0;
                s->ref_span = L + b;
            }
        } else {
            s->n_cigar = 1;
            s->cigar[0] = (L << 4) | 0;
            s->ref_span = L;
        }
        s->nm = (uint8_t)rbelow(&r, 4);
    }
    if (P->aux_mode == 1) {
        /* MM:Z:C+m?,d1,d2,...;  deltas ~ Geometric(0.05) over ~l_seq/4 C bases */
        uint64_t budget = L / 4, used = 0;
        uint32_t n = 0, bytes = 5; /* "C+m?" + ';' */
        s->mm_seed = rnext(&r);
        rng_t m = {s->mm_seed};
        for (;;) {
            double u = runif(&m);
            uint64_t d = (uint64_t)floor(log(1.0 - u) / log(1.0 - GEOM_P));
            if (used + d + 1 > budget) break;
            used += d + 1;
            char tmp[32];
            bytes += 1 + (uint32_t)snprintf(tmp, sizeof tmp, "%llu", (unsigned long long)d);
            ++n;
        }
        s->n_mm = n;
        s->mm_bytes = bytes + 1; /* NUL */
    }
}

static const char RG_VALUE[] = "rg1";

static uint64_t record_size(const synth_bam_params *P, const shape_t *s)
{
    uint64_t body = 32 + 14 /* name "syn%010llu\0" */ + 4ull * s->n_cigar + (s->l_seq + 1) / 2 + s->l_seq;
    if (P->aux_mode == 0) {
        body += 4 + 3 + sizeof RG_VALUE; /* NM:C:x, RG:Z:rg1\0 */
    } else {
        body += 3 + s->mm_bytes;                 /* MM:Z */
        body += 3 + 1 + 4 + s->n_mm;             /* ML:B:C */
        body += 3 + 1 + 4 + 2ull * s->l_seq;     /* fi:B:S kinetics */
    }
    return 4 + body;
}

static void write_record(const synth_bam_params *P, uint64_t i, const shape_t *s, int32_t pos,
                         uint8_t *o)
{
    rng_t r = {splitmix(P->seed * 0xC2B2AE3D27D4EB4Full ^ (i * 0x165667B19E3779F9ull) ^ 0xA5A5)};
    uint64_t size = record_size(P, s);
    w32(o, (uint32_t)(size - 4));
    w32(o + 4, (uint32_t)(P->unaligned ? -1 : 0));
    w32(o + 8, (uint32_t)pos);
    o[12] = 14;
    o[13] = P->unaligned ? 255 : 60;
    w16(o + 14, (uint16_t)(P->unaligned ? 4680 : reg2bin(pos, pos + (int64_t)s->ref_span)));
    w16(o + 16, (uint16_t)s->n_cigar);
    w16(o + 18, s->flag);
    w32(o + 20, s->l_seq);
    w32(o + 24, (uint32_t)-1);
    w32(o + 28, (uint32_t)-1);
    w32(o + 32, 0);
    uint8_t *p = o + 36;
    char name[16];
    snprintf(name, sizeof name, "syn%010llu", (unsigned long long)(i % 10000000000ull));
    memcpy(p, name, 13);
    p[13] = 0;
    p += 14;
    for (uint32_t k = 0; k < s->n_cigar; ++k, p += 4) w32(p, s->cigar[k]);
    /* 4-bit seq: A=1 C=2 G=4 T=8, high nibble first */
    static const uint8_t NIB[4] = {1, 2, 4, 8};
    uint32_t L = s->l_seq;
    uint64_t bitsrc = 0;
    int avail = 0;
    for (uint32_t k = 0; k < L; k += 2) {
        if (avail < 4) { bitsrc = rnext(&r); avail = 64; }
        uint8_t hi = NIB[bitsrc & 3];
        uint8_t lo = (k + 1 < L) ? NIB[(bitsrc >> 2) & 3] : 0;
        bitsrc >>= 4;
        avail -= 4;
        *p++ = (uint8_t)((hi << 4) | lo);
    }
    for (uint32_t k = 0; k < L; k += 5) {
        uint64_t v = rnext(&r);
        for (uint32_t j = 0; j < 5 && k + j < L; ++j) { p[k + j] = QTAB[v & 4095]; v >>= 12; }
    }
    p += L;
    if (P->aux_mode == 0) {
        p[0] = 'N'; p[1] = 'M'; p[2] = 'C'; p[3] = s->nm; p += 4;
        p[0] = 'R'; p[1] = 'G'; p[2] = 'Z'; p += 3;
        memcpy(p, RG_VALUE, sizeof RG_VALUE); p += sizeof RG_VALUE;
    } else {
        p[0] = 'M'; p[1] = 'M'; p[2] = 'Z'; p += 3;
        memcpy(p, "C+m?", 4); p += 4;
        rng_t m = {s->mm_seed};
        for (uint32_t k = 0; k < s->n_mm; ++k) {
            double u = runif(&m);
            uint64_t d = (uint64_t)floor(log(1.0 - u) / log(1.0 - GEOM_P));
            *p++ = ',';
            p += sprintf((char *)p, "%llu", (unsigned long long)d);
        }
        *p++ = ';';
        *p++ = 0;
        p[0] = 'M'; p[1] = 'L'; p[2] = 'B'; p[3] = 'C'; w32(p + 4, s->n_mm); p += 8;
        for (uint32_t k = 0; k < s->n_mm; ++k) *p++ = (uint8_t)rnext(&r);
        p[0] = 'f'; p[1] = 'i'; p[2] = 'B'; p[3] = 'S'; w32(p + 4, L); p += 8;
        for (uint32_t k = 0; k < L; k += 4) {
            uint64_t v = rnext(&r);
            for (uint32_t j = 0; j < 4 && k + j < L; ++j) { w16(p + 2 * (k + j), (uint16_t)v); v >>= 16; }
        }
        p += 2ull * L;
    }
    if ((uint64_t)(p - o) != size) {
        fprintf(stderr, "synth: record %llu size mismatch\n", (unsigned long long)i);
        abort();
    }
}

typedef struct {
    const synth_bam_params *P;
    shape_t *shapes;
    int32_t *pos;
    uint64_t *offs;
    uint8_t *u;
    uint64_t lo, hi;
} gen_job_t;

static void *gen_shapes(void *a)
{
    gen_job_t *j = (gen_job_t *)a;
    for (uint64_t i = j->lo; i < j->hi; ++i) make_shape(j->P, i, &j->shapes[i]);
    return NULL;
}

static void *gen_records(void *a)
{
    gen_job_t *j = (gen_job_t *)a;
    for (uint64_t i = j->lo; i < j->hi; ++i)
        write_record(j->P, i, &j->shapes[i], j->pos[i], j->u + j->offs[i]);
    return NULL;
}

/* ---- BGZF framing ---- */

static const uint8_t EOF_MEMBER[28] = {0x1f, 0x8b, 0x08, 0x04, 0x00, 0x00, 0x00, 0x00, 0x00, 0xff,
                                       0x06, 0x00, 0x42, 0x43, 0x02, 0x00, 0x1b, 0x00, 0x03, 0x00,
                                       0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00, 0x00};

typedef struct {
    const uint8_t *u;
    uint64_t u_len;
    uint32_t cap;
    uint64_t n_members;
    uint8_t *cbuf;     /* n_members * 65536 scratch */
    uint32_t *csize;   /* whole member size */
    const synth_bgzf_params *B;
    int tid, nthreads;
    int failed;
} bgzf_job_t;

static int member_level(const synth_bgzf_params *B, uint64_t m)
{
    if (B->level_mode == 0) return B->level;
    uint64_t h = splitmix(B->seed ^ (m * 0x9FB21C651E98DF25ull));
    double u = (h >> 11) * (1.0 / 9007199254740992.0);
    if (u < B->stored_frac) return 0;
    static const int L[3] = {1, 6, 9};
    return L[(splitmix(h) >> 32) % 3];
}

static void *compress_members(void *a)
{
    bgzf_job_t *j = (bgzf_job_t *)a;
    z_stream zs;
    for (uint64_t m = (uint64_t)j->tid; m < j->n_members; m += (uint64_t)j->nthreads) {
        uint64_t off = m * (uint64_t)j->cap;
        uint32_t len = (uint32_t)((j->u_len - off) < j->cap ? (j->u_len - off) : j->cap);
        uint8_t *dst = j->cbuf + m * 65536ull;
        /* As noodles' deflate::encode (crates/fritillaria-bgzf/src/deflate.rs:58-95):
           a payload that compresses past 65,495 + 15 bytes is re-encoded at level 0. */
        int levels[2] = {member_level(j->B, m), 0};
        uint32_t clen = 0;
        int ok = 0;
        for (int li = 0; li < 2 && !ok; ++li) {
            memset(&zs, 0, sizeof zs);
            if (deflateInit2(&zs, levels[li], Z_DEFLATED, -15, 8,
                             j->B->strategy ? j->B->strategy : Z_DEFAULT_STRATEGY) != Z_OK) {
                j->failed = 1;
                return NULL;
            }
            zs.next_in = (Bytef *)(j->u + off);
            zs.avail_in = len;
            zs.next_out = dst + 18;
            zs.avail_out = 65536 - 26;
            int r = deflate(&zs, Z_FINISH);
            clen = (uint32_t)zs.total_out;
            deflateEnd(&zs);
            ok = r == Z_STREAM_END && clen <= 65495 + 15;
        }
        if (!ok) { j->failed = 1; return NULL; }
        uint32_t bsize = 18 + clen + 8;
        static const uint8_t H[16] = {0x1f, 0x8b, 0x08, 0x04, 0, 0, 0, 0, 0, 0xff, 0x06, 0x00, 'B', 'C', 0x02, 0x00};
        memcpy(dst, H, 16);
        w16(dst + 16, (uint16_t)(bsize - 1));
        w32(dst + 18 + clen, (uint32_t)crc32(0L, j->u + off, len));
        w32(dst + 18 + clen + 4, len);
        j->csize[m] = bsize;
    }
    return NULL;
}

int synth_bgzf(const uint8_t *u, uint64_t u_len, const synth_bgzf_params *B, uint8_t **out,
               uint64_t *out_len)
{
    uint32_t cap = B->payload_cap ? B->payload_cap : 65495;
    uint64_t n = (u_len + cap - 1) / cap;
    int T = B->threads > 0 ? B->threads : 1;
    uint8_t *cbuf = (uint8_t *)malloc((n ? n : 1) * 65536ull);
    uint32_t *csize = (uint32_t *)calloc(n ? n : 1, sizeof(uint32_t));
    if (!cbuf || !csize) { free(cbuf); free(csize); return -1; }
    bgzf_job_t *jobs = (bgzf_job_t *)calloc((size_t)T, sizeof(bgzf_job_t));
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    for (int t = 0; t < T; ++t) {
        jobs[t] = (bgzf_job_t){u, u_len, cap, n, cbuf, csize, B, t, T, 0};
        pthread_create(&th[t], NULL, compress_members, &jobs[t]);
    }
    int failed = 0;
    for (int t = 0; t < T; ++t) { pthread_join(th[t], NULL); failed |= jobs[t].failed; }
    free(jobs);
    free(th);
    if (failed) { free(cbuf); free(csize); return -2; }
    uint64_t total = 28;
    for (uint64_t m = 0; m < n; ++m) total += csize[m];
    uint8_t *o = (uint8_t *)malloc(total);
    uint64_t p = 0;
    for (uint64_t m = 0; m < n; ++m) { memcpy(o + p, cbuf + m * 65536ull, csize[m]); p += csize[m]; }
    if (B->eof_member) { memcpy(o + p, EOF_MEMBER, 28); p += 28; }
    free(cbuf);
    free(csize);
    *out = o;
    *out_len = p;
    return 0;
}

int synth_bam_stream(const synth_bam_params *P, uint8_t **out, uint64_t *out_len, uint64_t *records_start)
{
    pthread_once(&qonce, build_qtab);
    uint64_t n = P->n_records;
    int T = P->threads > 0 ? P->threads : 1;
    shape_t *shapes = (shape_t *)malloc(sizeof(shape_t) * (n ? n : 1));
    int32_t *pos = (int32_t *)malloc(sizeof(int32_t) * (n ? n : 1));
    uint64_t *offs = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    if (!shapes || !pos || !offs) return -1;
    pthread_t *th = (pthread_t *)calloc((size_t)T, sizeof(pthread_t));
    gen_job_t *jobs = (gen_job_t *)calloc((size_t)T, sizeof(gen_job_t));
    for (int t = 0; t < T; ++t) {
        jobs[t] = (gen_job_t){P, shapes, pos, offs, NULL, n * (uint64_t)t / T, n * (uint64_t)(t + 1) / T};
        pthread_create(&th[t], NULL, gen_shapes, &jobs[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);

    /* Sorted uniform positions: normalised cumulative exponential gaps are the
       order statistics of n uniforms. */
    if (P->unaligned) {
        for (uint64_t i = 0; i < n; ++i) pos[i] = -1;
    } else {
        rng_t r = {splitmix(P->seed ^ 0x5050505050ull)};
        double acc = 0;
        double *g = (double *)malloc(sizeof(double) * (n + 1));
        for (uint64_t i = 0; i <= n; ++i) {
            double u = runif(&r);
            acc += -log(1.0 - u);
            g[i] = acc;
        }
        double span = (double)(P->ref_len > 200 ? P->ref_len - 200 : 1);
        for (uint64_t i = 0; i < n; ++i) pos[i] = (int32_t)floor(g[i] / acc * span);
        free(g);
    }

    /* header */
    char text[256];
    int tl = snprintf(text, sizeof text,
                      "@HD\tVN:1.6\tSO:%s\n@SQ\tSN:chr1\tLN:%u\n@RG\tID:rg1\tSM:synthetic\n",
                      P->unaligned ? "unknown" : "coordinate", P->ref_len);
    uint64_t hdr = 4 + 4 + (uint64_t)tl + 4 + 4 + 5 + 4;
    offs[0] = hdr;
    for (uint64_t i = 0; i < n; ++i) offs[i + 1] = offs[i] + record_size(P, &shapes[i]);
    uint64_t total = offs[n];
    uint8_t *u = (uint8_t *)malloc(total ? total : 1);
    if (!u) return -1;
    memcpy(u, "BAM\1", 4);
    w32(u + 4, (uint32_t)tl);
    memcpy(u + 8, text, (size_t)tl);
    uint8_t *h = u + 8 + tl;
    w32(h, 1);
    w32(h + 4, 5);
    memcpy(h + 8, "chr1", 5);
    w32(h + 13, P->ref_len);
    for (int t = 0; t < T; ++t) {
        jobs[t].u = u;
        pthread_create(&th[t], NULL, gen_records, &jobs[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
    free(th);
    free(jobs);
    free(shapes);
    free(pos);
    free(offs);
    *out = u;
    *out_len = total;
    *records_start = hdr;
    return 0;
}

void synth_free(void *p) { free(p); }
