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
    const uint8_t *lut_mask;/* len bytes, nonzero = residue seeds no word; or NULL */
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
/* Property check of HSP records against the sequences (extent, context,
 * diagonal, recomputed score, cutoff, (oid, s_off, q_off) order); returns
 * the number of bad records, *first_bad the first one's index or n. */
uint64_t hbo_hsp_check(const hbo_query *q, const uint8_t *residues, const uint64_t *offsets,
                       uint32_t n_subjects, const hbo_hsp *h, uint64_t n, uint64_t *first_bad);

/* --- preliminary gapped extension (hbo_gap.c) ---------------------------- */
typedef struct hbo_gap_params {
    int32_t gap_open;   /* 11 for BLOSUM62 */
    int32_t gap_extend; /* 1 */
    int32_t x_dropoff;  /* raw gapped X-drop (38 = 15 bits with lambda 0.267) */
} hbo_gap_params;

typedef struct hbo_ghsp {
    uint32_t oid;
    int32_t ctx;
    int32_t q_start, q_end;  /* context-relative query range, end exclusive */
    int32_t s_start, s_end;  /* subject range, end exclusive */
    int32_t q_gapped_start;  /* start point of the extension (context-relative) */
    int32_t s_gapped_start;
    int32_t score;           /* gapped raw score */
    int32_t pad;
} hbo_ghsp;

/* Score-only X-drop DP from the origin of A[1..M] x B[1..N] (reverse = 1:
 * A[M-1..0] x B[N-1..0]).  Returns the best score; *a_off / *b_off receive the
 * letters of A and B consumed by the best-scoring alignment. */
int32_t hbo_semigapped(const int8_t *m, const uint8_t *A, const uint8_t *B, int32_t M,
                       int32_t N, int reverse, int32_t gap_open, int32_t gap_extend,
                       int32_t x_dropoff, int32_t *a_off, int32_t *b_off);
/* Start point of a gapped extension inside an ungapped HSP (query offset). */
int32_t hbo_gapped_start(const int8_t *m, const uint8_t *q, const uint8_t *s, int32_t q_start,
                         int32_t q_len, int32_t s_start, int32_t s_len);
/* Two-sided gapped extension through (q_off, s_off).  coords receives
 * q_start, q_end, s_start, s_end; returns the score. */
int32_t hbo_prot_gapped(const int8_t *m, const uint8_t *q, int32_t q_len, const uint8_t *s,
                        int32_t s_len, int32_t q_off, int32_t s_off,
                        const hbo_gap_params *gp, int32_t *coords);
/* One subject's ungapped HSPs (any order) -> its gapped HSPs in save order;
 * *extensions counts the extensions performed (HSPs not contained). */
int hbo_gapped_subject(const hbo_query *q, const uint8_t *subj, int32_t slen, const hbo_hsp *in,
                       int32_t n_in, const hbo_gap_params *gp, const int32_t *gap_cutoff,
                       hbo_ghsp *out, int32_t *n_out, int32_t *extensions);
/* Whole database: in is grouped by ascending oid (as hbo_scan_db returns
 * it); counts receives 2 ints per subject (gapped HSPs saved, extensions).
 * Returns the number of gapped HSPs, ordered by (oid, save order); *out is
 * released with hbo_free. */
int64_t hbo_gapped_db(const hbo_query *q, const uint8_t *residues, const uint64_t *offsets,
                      uint32_t n_subjects, const hbo_hsp *in, uint64_t n_in,
                      const hbo_gap_params *gp, const int32_t *gap_cutoff, int n_threads,
                      hbo_ghsp **out, int32_t *counts);

/* --- H-BLAST GPU database volume (.hDb) and index (.hDbI), hbo_hdb.c --- */
typedef struct hbo_hdb_volume {
    const char *name;         /* .hDb file name */
    uint64_t bytes;           /* .hDb size */
    int32_t volume_seqs;      /* sequences in the volume */
    int32_t gpu_seqs;         /* sequences in the .hDb (the shortest ones) */
    int32_t store_limit;      /* oid after the last one in the .hDb */
    int32_t break_limit;      /* oid after the volume's last one */
    int32_t n_groups;
    const int32_t *group_len; /* n_groups entries */
    uint64_t cpu_rest_bytes;  /* residues of the sequences left to the CPU */
} hbo_hdb_volume;

/* Sequences of a volume that go to the .hDb. */
int32_t hbo_hdb_gpu_seqs(int32_t volume_seqs);
/* Group lengths (ceil(n_seqs / (blocksx*threadsx)) entries) and file size
 * of a .hDb holding sequences 0..n_seqs-1 (ascending length). */
int hbo_hdb_layout(const uint64_t *offsets, int32_t n_seqs, int32_t blocksx, int32_t threadsx,
                   int32_t *group_len, uint64_t *bytes);
/* Writes the .hDb bytes (out_bytes from hbo_hdb_layout). */
int hbo_hdb_pack(const uint8_t *residues, const uint64_t *offsets, int32_t n_seqs,
                 int32_t blocksx, int32_t threadsx, const int32_t *group_len, uint8_t *out,
                 uint64_t out_bytes);
/* Reads a .hDb back: offsets gets n_seqs + 1 entries; residues (NULL to
 * size) the sequences, each ending after its last non-zero byte. */
int hbo_hdb_unpack(const uint8_t *hdb, uint64_t bytes, int32_t n_seqs, int32_t blocksx,
                   int32_t threadsx, const int32_t *group_len, uint8_t *residues,
                   uint64_t *offsets);
/* The .hDbI text; returns its length (text NULL or cap too small: only
 * the length). */
int64_t hbo_hdbi_format(int32_t blocksx, int32_t threadsx, int32_t percentage, int32_t n_volumes,
                        const hbo_hdb_volume *vols, char *text, uint64_t cap);

/* --- CPU/GPU split of a volume (hbo_balance.c) --------------------------- */
typedef struct hbo_options {  /* the H-BLAST_runtime_options file */
    float gpu_ratio;
    int32_t fixed_ratio;
    double capability_rate;
    int32_t max_hits;
    int32_t mem_ecc;
    int32_t load_balancer;
    int32_t load_factor;
    int32_t subject_chunk;
    int32_t base_value;
} hbo_options;

void hbo_options_default(hbo_options *o);
/* 0, or HBO_ERR_ARG when the text has fewer than 18 lines. */
int hbo_options_parse(const char *text, uint64_t len, hbo_options *o);
/* GPU share of a volume at *ratio_io (-1: by cpu_threads); group_size =
 * blocksx * threadsx. */
int hbo_split(const int32_t *glen, int32_t ngroups, int32_t nseq, uint64_t file_bytes,
              int32_t threadsx, int32_t group_size, int32_t cpu_threads, float *ratio_io,
              int32_t *fixed_io, int32_t *out_seqs, int32_t *out_groups, uint64_t *out_bytes);
/* Ratio for volume v + 1 (rings as the product's hb_balance_in). */
int hbo_next_ratio(const hbo_options *o, unsigned int v, float ratio, unsigned int cpu_threads,
                   unsigned int gpu_cards, const uint64_t *vol_bytes, const float *ratio_ring,
                   const int64_t *hsp_ring, int gapped, int query_length,
                   unsigned int avg_query_length, double gpu_t, double cpu_t, float *out);

/* --- SEG low-complexity masking (hbo_seg.c) ------------------------------- */
/* Shannon entropy (bits) of counts[28] over the 20 standard letters (the
 * other codes must be 0), and ln P of that composition over n letters. */
double hbo_seg_entropy(const int32_t *counts);
double hbo_seg_lnprob(const int32_t *counts, int32_t n);
int hbo_seg_mask(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                 const int32_t *ctx_len, int32_t window, double locut, double hicut,
                 int32_t maxbogus, int32_t maxtrim, uint8_t *mask);

#define HBO_ERR_ARG (-1)
#define HBO_ERR_NOMEM (-2)
#define HBO_ERR_RESIDUE (-3)

#ifdef __cplusplus
}
#endif
#endif
