/* synth.h -- synthetic BAM streams and BGZF framing for tests and benchmarks. */
#ifndef FR_SYNTH_H
#define FR_SYNTH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t n_records;
    uint64_t seed;
    int32_t read_mode;   /* 0 fixed length, 1 uniform [len_min, len_max], 2 log-normal */
    uint32_t len_fixed, len_min, len_max;
    double ln_median, ln_sigma;
    uint32_t len_cap;
    double indel_p;      /* probability of one insertion or deletion per read */
    int32_t aux_mode;    /* 0: NM:C + RG:Z; 1: MM:Z + ML:B:C + fi:B:S kinetics */
    int32_t unaligned;   /* refID -1, pos -1, flag 4, no CIGAR */
    uint32_t ref_len;
    int32_t threads;
    uint32_t cap_every;  /* log-normal lengths: records k*cap_every + cap_every/2 take len_cap (0: none) */
} synth_bam_params;

typedef struct {
    uint64_t seed;
    int32_t level_mode;  /* 0: every member at `level`; 1: uniform {1,6,9}, `stored_frac` at 0 */
    int32_t level;
    double stored_frac;
    uint32_t payload_cap; /* uncompressed bytes per member; noodles uses 65,495 */
    int32_t strategy;     /* zlib strategy, 0 = default */
    int32_t eof_member;
    int32_t threads;
} synth_bgzf_params;

int synth_bam_stream(const synth_bam_params *P, uint8_t **out, uint64_t *out_len,
                     uint64_t *records_start);
int synth_bgzf(const uint8_t *u, uint64_t u_len, const synth_bgzf_params *B, uint8_t **out,
               uint64_t *out_len);
void synth_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
