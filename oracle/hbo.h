/*
 * hbo.h -- CPU oracle for the BLASTP/BLASTX seed-extension stage.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product library (libhbseed.so) never
 * links, loads or calls it.
 *
 * This is a sequential restatement of the two-hit ungapped word finder that
 * H-BLAST offloads to the GPU and that NCBI BLAST+ 2.2.28 runs on the CPU
 * (BlastAaWordFinder, two-hit path, aa_ungapped.c -- not vendored in the
 * reference; see hbo.c for the reference lines each step follows).
 *
 * Conventions (shared with include/hb_seed.h):
 *   - residues are ncbistdaa codes 0..27; code 0 is the gap/sentinel letter;
 *   - the query is a concatenation of contexts separated by code-0 sentinels;
 *   - word size 3; a 3-mer indexes cell (a<<10)|(b<<5)|c  (32768 cells);
 *   - scores are int32, the matrix is int8.
 */
#ifndef HBO_H
#define HBO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBO_ALPHA 28
#define HBO_NCELLS 32768

typedef struct hbo_hsp {
    uint32_t oid;     /* subject index */
    int32_t ctx;      /* query context of q_off */
    int32_t q_start;  /* HSP start in the concatenated query */
    int32_t s_start;  /* HSP start in the subject */
    int32_t q_off;    /* query offset of the hit that triggered the extension */
    int32_t s_off;    /* subject offset of that hit */
    int32_t len;      /* HSP length */
    int32_t score;    /* ungapped score */
} hbo_hsp;

typedef struct hbo_query {
    const uint8_t *seq;     /* concatenated query, len bytes */
    int32_t len;
    int32_t n_ctx;
    const int32_t *ctx_off; /* n_ctx ascending start offsets */
    const int32_t *ctx_len; /* n_ctx lengths */
    const int32_t *xdrop;   /* n_ctx raw X-drop values */
    const int32_t *cutoff;  /* n_ctx raw cutoff scores */
    int32_t word_size;      /* must be 3 */
    int32_t threshold;      /* neighbourhood threshold T */
    int32_t window;         /* two-hit window */
    const int8_t *matrix;   /* 28x28 row-major, or NULL for BLOSUM62 */
} hbo_query;

typedef struct hbo_lut {
    uint32_t *start;        /* HBO_NCELLS + 1 cell starts (CSR) */
    uint32_t *entries;      /* query offsets, ascending within a cell */
    uint32_t n_entries;
} hbo_lut;

/* 28x28 BLOSUM62 in ncbistdaa order (row-major). */
const int8_t *hbo_blosum62(void);

/* Neighbourhood lookup table.  Returns 0 or a negative error. */
int hbo_lut_build(const hbo_query *q, hbo_lut *out);
void hbo_lut_free(hbo_lut *lut);

/* Ungapped extensions (exported for the pinning tests). */
int32_t hbo_extend_left(const int8_t *m, const uint8_t *s, const uint8_t *q,
                        int32_t s_off, int32_t q_off, int32_t dropoff,
                        int32_t maxscore, int32_t *length);
int32_t hbo_extend_right(const int8_t *m, const uint8_t *s, int32_t slen,
                         const uint8_t *q, int32_t qlen, int32_t s_off,
                         int32_t q_off, int32_t dropoff, int32_t maxscore,
                         int32_t *length, int32_t *s_last_off);
int32_t hbo_extend_two_hit(const int8_t *m, const uint8_t *s, int32_t slen,
                           const uint8_t *q, int32_t qlen, int32_t s_left_off,
                           int32_t s_right_off, int32_t q_right_off,
                           int32_t dropoff, int32_t *hsp_q, int32_t *hsp_s,
                           int32_t *hsp_len, int32_t *right_extend,
                           int32_t *s_last_off);

/* Search subjects [0, n_subjects) of a packed database (subject i is
 * residues[offsets[i] .. offsets[i+1])).  HSPs come back ordered by
 * (oid, discovery order), which is (oid, s_off, q_off).  counters, if not
 * NULL, receives 3 ints per subject: lookup hits, extensions, HSPs.
 * Returns the HSP count (>= 0) or a negative error; *out is malloc'd and
 * released with hbo_free. */
int64_t hbo_scan_db(const hbo_query *q, const hbo_lut *lut,
                    const uint8_t *residues, const uint64_t *offsets,
                    uint32_t n_subjects, int n_threads, hbo_hsp **out,
                    uint32_t *counters);
void hbo_free(void *p);

#define HBO_ERR_ARG (-1)
#define HBO_ERR_NOMEM (-2)
#define HBO_ERR_RESIDUE (-3)

#ifdef __cplusplus
}
#endif
#endif
