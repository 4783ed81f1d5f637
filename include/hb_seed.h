/*
 * hb_seed.h -- C ABI of the MI355X seed-extension stage (BLOSUM62 word scan,
 * two-hit diagonal tracking, X-drop ungapped extension) for BLASTP and
 * six-frame BLASTX queries.
 *
 * Drop-in boundary: these entry points replace the extern "C" API H-BLAST
 * exports to NCBI's engine (h_blast_v1.1/H-BLAST/ncbi_blast_files/
 * gpu_cpu_common.h:43-126, implemented in H-BLAST/h_blast.cu).  The
 * reference passes NCBI structs (LookupTableWrap, Blast_ExtendWord,
 * BlastInitialWordParameters, BLAST_SequenceBlk, BlastQueryInfo); this ABI
 * takes the plain arrays those structs hold, so a caller without NCBI
 * headers can use it.  All functions return an int status (HB_OK or a
 * negative HB_ERR_*) instead of calling exit() (H-BLAST/h_blast.cu:311-332).
 *
 * Conventions
 *   - residues: ncbistdaa codes 0..27 (0 = gap/sentinel, 1 = A, ... 27 = J);
 *   - query: contexts concatenated, separated by code-0 sentinels; offsets
 *     in results are positions in that concatenation (as NCBI's
 *     BlastInitHitList holds them);
 *   - word size 3 only; scores are int32.
 *
 * Threading: an hb_ctx serialises its own calls; distinct contexts may be
 * used from distinct threads.  Handles are owned by the library and released
 * with the matching *_free; output arrays are owned by the caller.
 */
#ifndef HB_SEED_H
#define HB_SEED_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HB_OK 0
#define HB_ERR_ARG (-1)          /* invalid argument */
#define HB_ERR_NOMEM (-2)        /* host or device allocation failed */
#define HB_ERR_DEVICE (-3)       /* HIP runtime error */
#define HB_ERR_CAPACITY (-4)     /* caller's output array too small */
#define HB_ERR_RESIDUE (-5)      /* residue code >= 28 */
#define HB_ERR_STATE (-6)        /* handle used in the wrong state */
#define HB_ERR_UNSUPPORTED (-7)  /* e.g. word size != 3 */
#define HB_ERR_BUSY (-8)         /* serving queue full / not admitted */
#define HB_ERR_FORMAT (-9)       /* malformed .hDb volume or .hDbI index */

typedef struct hb_ctx hb_ctx;
typedef struct hb_db hb_db;
typedef struct hb_query hb_query;
typedef struct hb_scan hb_scan;

/* One initial HSP.  Replaces GPUBlastInitHitList
 * (ncbi_blast_files/gpu_cpu_common.h:31-40): full-width coordinates, the
 * context and the triggering hit's offsets (what BlastSaveInitHsp receives,
 * ncbi_blast_files/blast_engine.c:568-576). */
typedef struct hb_hsp {
    uint32_t oid;     /* subject index in the uploaded database */
    int32_t context;  /* query context (frame for BLASTX) */
    int32_t q_start;  /* HSP start, concatenated-query coordinates */
    int32_t s_start;  /* HSP start in the subject */
    int32_t q_off;    /* query offset of the second (triggering) hit */
    int32_t s_off;    /* subject offset of that hit */
    int32_t len;      /* HSP length */
    int32_t score;    /* ungapped raw score */
} hb_hsp;

/* Per-subject counters.  Replaces the d_Hits[4*oid .. 4*oid+3] quadruple
 * {lookup hits, HSPs, extensions, list base | -1}
 * (H-BLAST/h_blast_kernel.cu:462-492, consumed at
 * ncbi_blast_files/blast_engine.c:540-545).  There is no "-1 = recompute on
 * the CPU" state: the device never drops HSPs. */
typedef struct hb_subject_stats {
    uint32_t lookup_hits;  /* word hits (sum of cell sizes over positions) */
    uint32_t extensions;   /* two-hit extensions performed */
    uint32_t n_hsp;        /* HSPs with score >= cutoff */
    int64_t first;         /* index of the subject's first HSP in the list, -1 if none */
} hb_subject_stats;

/* Search parameters.  The reference reads these from NCBI structs
 * (lookup->word_length, ewp->diag_table->window, the LUT threshold;
 * H-BLAST/h_blast.cu:959-983). */
typedef struct hb_params {
    int32_t word_size;     /* must be 3 */
    int32_t threshold;     /* neighbourhood threshold T (default 11) */
    int32_t window;        /* two-hit window (default 40) */
    const int8_t *matrix;  /* 28x28 row-major scores, or NULL for BLOSUM62 */
} hb_params;

/* Fills the defaults: word 3, T 11, window 40, BLOSUM62. */
void hb_params_default(hb_params *p);
/* Default per-context X-drop (16) and cutoff (41): BLOSUM62 ungapped
 * lambda 0.3176, K 0.134; X = ceil(7 ln2 / lambda); cutoff =
 * floor((22 ln2 + ln K) / lambda). */
#define HB_DEFAULT_XDROP 16
#define HB_DEFAULT_CUTOFF 41

const char *hb_strerror(int err);
const int8_t *hb_blosum62(void);  /* 28x28 row-major */

/* --- device context ----------------------------------------------------
 * Replaces GPU_BLAST_check_availability / SetNumOfGpuCards /
 * set_first_gpu_id (ncbi_blast_files/gpu_cpu_common.h:99,111,79). */
int hb_create(int device, hb_ctx **out);
int hb_destroy(hb_ctx *ctx);
/* Free / total device memory in bytes (the served form's admission and
 * hb_check_memory's default budget use it; hb_check_memory is the
 * counterpart of H_BLAST_check_memory, gpu_cpu_common.h:93-102). */
int hb_mem_info(hb_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);
int hb_synchronize(hb_ctx *ctx);
/* Page-locked host memory (hipHostMalloc) for the arrays a caller hands to
 * hb_db_upload / hb_db_upload_hdb: reading a database file straight into it
 * lets the upload run at the host link's full rate instead of through the
 * runtime's pageable staging (the reference reads volumes into ordinary
 * memory and copies per call, h_blast.cu:587-653, 1094).  *out is NULL on
 * failure; release with hb_host_free. */
int hb_host_alloc(hb_ctx *ctx, uint64_t bytes, void **out);
int hb_host_free(hb_ctx *ctx, void *p);

/* --- packed subject database -------------------------------------------
 * Replaces Read_H_BLAST_GPU_DB + the per-call DB upload
 * (H-BLAST/h_blast.cu:587-653, 1094-1096): the database is uploaded once
 * and stays resident in HBM across scans (its residues twice: in oid order,
 * and in the scan's length-descending order).  Subject i is
 * residues[offsets[i] .. offsets[i+1]); offsets has n_subjects+1 entries. */
int hb_db_upload(hb_ctx *ctx, const uint8_t *residues, const uint64_t *offsets,
                 uint32_t n_subjects, hb_db **out);
int hb_db_free(hb_db *db);
uint32_t hb_db_subjects(const hb_db *db);
uint64_t hb_db_residues(const hb_db *db);
/* Copies the resident database back: residues (res_cap bytes, at least
 * hb_db_residues()) and offsets (hb_db_subjects() + 1 entries); for tests
 * and for writing a loaded volume out again. */
int hb_db_fetch(const hb_db *db, uint8_t *residues, uint64_t res_cap, uint64_t *offsets);

/* --- query ---------------------------------------------------------------
 * Builds the neighbourhood lookup table, presence vector and query tables
 * on the device.  Replaces the LUT/PV/matrix/constant staging of H_BLAST
 * (H-BLAST/h_blast.cu:940-1091); the NCBI lookup table itself
 * (BlastAaLookupTable) is rebuilt here from the query.
 *   seq/len:      concatenated query (contexts separated by code 0)
 *   ctx_off/len:  n_ctx ascending context ranges inside seq
 *   xdrop/cutoff: per-context raw X-drop and cutoff score
 *                 (word_params->cutoffs[k], H-BLAST/h_blast.cu:968-975) */
int hb_query_prepare(hb_ctx *ctx, const uint8_t *seq, int32_t len, int32_t n_ctx,
                     const int32_t *ctx_off, const int32_t *ctx_len,
                     const int32_t *xdrop, const int32_t *cutoff,
                     const hb_params *params, hb_query **out);
/* As hb_query_prepare, with a lookup-table mask: lut_mask[i] != 0 keeps the
 * words that cover residue i out of the lookup table (NULL: no mask).  NCBI
 * builds its table over the unmasked segments of each context only
 * (lookup_segments: SEG / low-complexity filtering, the default for BLASTX,
 * or user masks); the reference copies that table (h_blast.cu:940-958).
 * Extensions still read the residues as given (soft masking); for hard
 * masking the caller replaces masked residues by X (21) as NCBI does. */
int hb_query_prepare_masked(hb_ctx *ctx, const uint8_t *seq, int32_t len, int32_t n_ctx,
                            const int32_t *ctx_off, const int32_t *ctx_len,
                            const int32_t *xdrop, const int32_t *cutoff,
                            const hb_params *params, const uint8_t *lut_mask,
                            hb_query **out);
int hb_query_free(hb_query *q);
/* Where a plain scan of q keeps its per-diagonal two-hit state (the
 * reference keeps it in global memory, ewp->diag_table, one array per
 * thread: h_blast.cu:959-966):
 *   HB_STATE_AUTO   - LDS (one 32-bit word per diagonal) unless LDS would
 *                     starve the CU of wavefronts or the database holds a
 *                     subject of 4,194,000 residues or more, else global
 *                     memory (the default);
 *   HB_STATE_LDS    - LDS (HB_ERR_UNSUPPORTED if the query's words do not
 *                     fit, or, at the scan, if the subjects are too long);
 *   HB_STATE_GLOBAL - a global-memory array per wavefront (32-bit words; the
 *                     ordering markers in an LDS table).
 * Results are identical; only speed differs.  Takes effect for later
 * hb_scan_start calls. */
#define HB_STATE_AUTO 0
#define HB_STATE_LDS 1
#define HB_STATE_GLOBAL 2
#define HB_STATE_LDS16 3   /* accepted for compatibility: the same as
                              HB_STATE_LDS (there is one LDS format) */
int hb_query_set_state_mode(hb_query *q, int mode);
/* The requested placement (HB_STATE_AUTO unless set; HB_STATE_LDS after
 * HB_STATE_LDS16 was set). */
int hb_query_state_mode(const hb_query *q);
/* Testing aid: the scan clears its diagonal state whenever the positions it
 * stores (subject offsets accumulated over the subjects a wavefront scans)
 * would pass a limit set by the placement; this lowers that limit to
 * `interval` residues (0: the placement's own limit) so that tests reach
 * the reset on small databases.  The interval must exceed twice the
 * longest subject plus the window (else the scan returns
 * HB_ERR_UNSUPPORTED).  Results are identical for every interval. */
int hb_query_set_reset_interval(hb_query *q, int32_t interval);
/* Number of (query offset, cell) entries of the lookup table. */
int hb_query_lut_info(const hb_query *q, uint32_t *n_entries, uint32_t *n_cells_used);
/* Copies the device lookup table back (CSR over 32768 cells:
 * cell_start[32769], entries[n_entries]); for tests. */
int hb_query_lut_fetch(const hb_query *q, uint32_t *cell_start, uint32_t *entries,
                       uint32_t cap);

/* Helpers that build the concatenated query of a BLASTP query (1 context)
 * or a six-frame translated BLASTX query (frames +1,+2,+3,-1,-2,-3, one
 * code-0 sentinel between frames, standard genetic code, stop = '*' (25),
 * ambiguous codon = X (21)).  nt is IUPAC text (ACGTN..., case-insensitive).
 * out_seq needs hb_blastx_concat_len(nt_len) bytes; ctx arrays need 6. */
int32_t hb_blastx_concat_len(int32_t nt_len);
int hb_blastx_translate(const char *nt, int32_t nt_len, uint8_t *out_seq,
                        int32_t *ctx_off, int32_t *ctx_len);
/* IUPAC protein letters <-> ncbistdaa codes. */
int hb_encode_protein(const char *text, int32_t len, uint8_t *out);
/* Protein FASTA text -> packed database arrays for hb_db_upload, the input
 * H-BLAST takes from NCBI's database (makeblastdb's FASTA reader, then
 * CreateH_BLAST_database, ncbi_blast_files/functions.cpp:785-1165).  Each
 * '>' line starts a subject; residue lines are encoded to ncbistdaa
 * (letters case-insensitive, '*' and '-' allowed; spaces, tabs, digits and
 * CR ignored); other characters give HB_ERR_RESIDUE with *n_res set to the
 * text offset of the first one.  Call with residues/offsets NULL to size:
 * *n_res and *n_seq receive the totals.  With buffers, residues holds
 * res_cap bytes, offsets seq_cap + 1 entries (subject i is
 * residues[offsets[i] .. offsets[i+1])), and the optional titles /
 * titles_len seq_cap entries (each '>' line's text offset after the '>' and
 * its length); HB_ERR_CAPACITY if they are too small (totals still set). */
int hb_fasta_parse(const char *text, uint64_t len, uint8_t *residues, uint64_t res_cap,
                   uint64_t *offsets, uint64_t *titles, uint32_t *titles_len,
                   uint32_t seq_cap, uint64_t *n_res, uint32_t *n_seq);

/* --- H-BLAST GPU database volumes (.hDb) and their index (.hDbI) --------
 * CreateH_BLAST_database (ncbi_blast_files/functions.cpp:785-1165) turns
 * each length-sorted NCBI volume into one .hDb file holding the volume's
 * shortest hb_hdb_gpu_seqs(n) sequences (the rest stay with NCBI's CPU
 * path): groups of blocksx * threadsx sequences, each group padded to its
 * longest member rounded up to 16 residues; inside a group, blocks of
 * threadsx sequences whose residues are interleaved in 4-byte stripes
 * (residues 4i..4i+3 of the block's t-th sequence at byte
 * (t + i * threadsx) * 4 of the block), zero-padded.  The file does not
 * store sequence lengths: a sequence ends after its last non-zero byte (a
 * trailing gap code 0 is not recovered).  The .hDbI text index
 * (Write_GPU_information_file*, functions.cpp:257-309) records the grid
 * and, per volume, the file, its size, counts and group lengths.  The
 * reference reads it with Read_GPU_DB_information_file*
 * (functions.cpp:312-446) and the file with Read_H_BLAST_GPU_DB
 * (H-BLAST/h_blast.cu:587-653). */
typedef struct hb_hdb_volume {
    const char *name;         /* .hDb file name; hb_hdbi_parse points into its text */
    uint32_t name_len;
    uint64_t bytes;           /* .hDb size */
    int32_t volume_seqs;      /* sequences in the volume */
    int32_t gpu_seqs;         /* sequences in the .hDb (the volume's first, shortest ones) */
    int32_t store_limit;      /* database oid after the last one in the .hDb */
    int32_t break_limit;      /* database oid after the volume's last one */
    int32_t n_groups;
    const int32_t *group_len; /* n_groups entries */
    uint64_t cpu_rest_bytes;  /* residues of the sequences left to the CPU */
} hb_hdb_volume;
typedef struct hb_hdb_index {
    int32_t blocksx, threadsx;  /* grid the volumes are laid out for */
    int32_t percentage;         /* as written (99) */
    int32_t n_volumes;
} hb_hdb_index;

/* Sequences of an n-sequence volume that go to its .hDb:
 * floor(floor(99.9% of n) / 32) * 32 (functions.cpp:842-843). */
int32_t hb_hdb_gpu_seqs(int32_t volume_seqs);
/* Layout of a .hDb holding sequences 0..n_seqs-1 of the packed arrays
 * (ascending length, else HB_ERR_ARG; functions.cpp:844-911): *n_groups =
 * ceil(n_seqs / (blocksx * threadsx)), group_len receives that many entries
 * (group_cap of room; NULL to size), *bytes the file size. */
int hb_hdb_layout(const uint64_t *offsets, int32_t n_seqs, int32_t blocksx, int32_t threadsx,
                  int32_t *group_len, int32_t group_cap, int32_t *n_groups, uint64_t *bytes);
/* Writes the .hDb bytes of that layout (functions.cpp:1009-1073); out_bytes
 * must equal the layout's size. */
int hb_hdb_pack(const uint8_t *residues, const uint64_t *offsets, int32_t n_seqs,
                int32_t blocksx, int32_t threadsx, const int32_t *group_len, int32_t n_groups,
                uint8_t *out, uint64_t out_bytes);
/* Writes the .hDbI text (names are NUL-terminated or name_len long); *len
 * receives its length.  text NULL sizes; HB_ERR_CAPACITY if cap is short. */
int hb_hdbi_format(const hb_hdb_index *index, const hb_hdb_volume *vols, char *text,
                   uint64_t cap, uint64_t *len);
/* Parses a .hDbI text.  index gets the header; with vols (vol_cap entries)
 * and group_len (group_cap entries) also the volumes, whose group_len
 * pointers point into group_len and names into text.  NULL arrays size:
 * *n_group_len receives the total of group lengths.  HB_ERR_FORMAT for a
 * malformed index, with *err_off the text offset of the offending line. */
int hb_hdbi_parse(const char *text, uint64_t len, hb_hdb_index *index, hb_hdb_volume *vols,
                  int32_t vol_cap, int32_t *group_len, uint64_t group_cap,
                  uint64_t *n_group_len, uint64_t *err_off);
/* Volume cut and length order of H-BLAST's makeblastdb -sort_volumes
 * (CWriteDB_Volume::WriteSequenceSort, ncbi_blast_files/
 * writedb_volume.cpp:255-513; makeblastdb.cpp:1054-1076): the n_seqs input
 * sequences (offsets, input order) fill volumes in order while their
 * sequence file stays within max_file_bytes (-max_file_sz; the file starts
 * at offset 1 and each sequence adds its length + 1, a volume's first
 * sequence always fits, a later one while offset + length < max_file_bytes,
 * writedb_files.hpp:435-494) and, when max_letters > 0, their letters
 * within max_letters; each volume is then ordered by ascending length as
 * the reference writes it (std::sort by descending length, written from the
 * back).  order[k] = the input index of the database's k-th sequence;
 * volume_seqs gets *n_volumes counts (NULL, or HB_ERR_CAPACITY past
 * volume_cap, to size).  The header and index files' limits depend on
 * deflines this library does not hold and are not applied. */
int hb_db_sort_volumes(const uint64_t *offsets, int32_t n_seqs, uint64_t max_file_bytes,
                       uint64_t max_letters, int32_t *order, int32_t *volume_seqs,
                       int32_t volume_cap, int32_t *n_volumes);
/* Writes an H-BLAST GPU database from length-sorted volumes, as
 * CreateH_BLAST_database does from NCBI's (functions.cpp:785-1165): the
 * database's k-th sequence is input sequence order[k] (a permutation, as
 * hb_db_sort_volumes gives it; NULL: the input order), volume_seqs[n_volumes]
 * counts the volumes' sequences (each volume must ascend in length up to its
 * .hDb share); files <prefix>.hDb (one volume) or <prefix>.NN.hDb and the
 * index <prefix>.hDbI. */
int hb_hdb_write(const char *prefix, const uint8_t *residues, const uint64_t *offsets,
                 int32_t n_seqs, const int32_t *order, const int32_t *volume_seqs,
                 int32_t n_volumes, int32_t blocksx, int32_t threadsx);
/* Uploads one .hDb volume (host bytes) and unpacks it on the device into a
 * resident database: subject i is the volume's i-th sequence (database oid
 * = the volume's first oid + i).  Replaces Read_H_BLAST_GPU_DB plus the
 * per-call upload (H-BLAST/h_blast.cu:587-653, 1094-1096).  HB_ERR_FORMAT
 * if bytes, group lengths (multiples of 16) and counts disagree. */
int hb_db_upload_hdb(hb_ctx *ctx, const uint8_t *hdb, uint64_t bytes, int32_t blocksx,
                     int32_t threadsx, int32_t n_seqs, const int32_t *group_len,
                     int32_t n_groups, hb_db **out);
/* Device time of the last hb_db_upload_hdb's unpacking kernels (lengths,
 * residues), in ms. */
int hb_db_hdb_info(const hb_db *db, float *lengths_ms, float *copy_ms);

/* --- CPU/GPU split of a volume (load balancing) --------------------------
 * H-BLAST searches a volume on both sides at once: the device scans a prefix
 * of the volume's .hDb while NCBI's CPU threads search the volume's other
 * sequences, and the prefix moves from volume to volume so that both sides
 * finish together (GPU_BLAST_PreliminarySearchEngine,
 * ncbi_blast_files/functions.cpp:1409-1520).  These three host routines make
 * that decision with the reference's arithmetic; the caller keeps the loop
 * (upload the prefix with hb_db_upload_hdb, hb_scan_start, its own CPU work,
 * hb_scan_wait, hb_scan_elapsed_ms). */

/* The H-BLAST_runtime_options file (Read_H_BLAST_runtime_options,
 * functions.cpp:142-203): a comment line before each of 9 values. */
typedef struct hb_runtime_options {
    float gpu_ratio;            /* GPU share of the workload, 0..1; -1 (no file):
                                   hb_hdb_split picks it from the CPU thread count */
    int32_t fixed_ratio;        /* 1: the ratio stays; 0: it adapts per volume */
    double capability_rate;     /* throughput of the GPU over one CPU thread (12) */
    int32_t max_hits_per_subject; /* the reference's per-subject HSP buffer (19; the
                                   device here keeps every HSP, so unused) */
    int32_t mem_ecc;            /* read, unused (0) */
    int32_t load_balancer;      /* 1: charge GPU HSPs to the CPU load estimate (1) */
    int32_t load_factor;        /* charge per HSP, ungapped (4) */
    int32_t subject_chunk;      /* CPU threads' chunk of subjects (200; the caller's) */
    int32_t base_value;         /* charge per HSP, gapped, for short queries (100) */
} hb_runtime_options;
/* The values in force when there is no options file (functions.cpp:14-26). */
void hb_runtime_options_default(hb_runtime_options *o);
/* Parses the file's text; values out of range take the reference's
 * fallbacks (ratio outside 0..1 -> 0, capability <= 0 -> 12, hits <= 0 ->
 * 20, negative factor / chunk / base -> 4 / 200 / 100).  HB_ERR_FORMAT if
 * the text has fewer than 18 lines (the reference then reads stale values). */
int hb_runtime_options_parse(const char *text, uint64_t len, hb_runtime_options *o);

/* The GPU's share of one volume at a workload ratio (get_GPU_DB_information,
 * functions.cpp:449-571).  Workload = group length x sequences.  With
 * ratio < 1 the GPU keeps the .hDb's first whole groups whose workload stays
 * within ratio x the total, then trunc(rest / (length x threadsx) x threadsx)
 * sequences of the next group: a prefix of the file that hb_db_upload_hdb
 * loads with out->gpu_seqs, the first out->n_groups group lengths and
 * out->bytes bytes.  *ratio == -1 picks the start ratio from cpu_threads
 * (1: 1, 2: 0.9, 3: 0.9, 4: 0.8, 5: 0.75, 6: 0.65, 7: 0.55, 8: 0.45, more:
 * 0.4; from 3 threads also *fixed = 0).  Unless *fixed == 1, *ratio
 * receives the ratio the split realises.  bytes is the .hDb size (kept when
 * nothing is cut).  Group byte totals use 64-bit arithmetic (the
 * reference's int products overflow past 2^31). */
typedef struct hb_hdb_share {
    int32_t gpu_seqs;   /* sequences the GPU scans: the .hDb's first ones */
    int32_t n_groups;   /* their groups (the last one may be partial) */
    uint64_t bytes;     /* the .hDb prefix holding them */
} hb_hdb_share;
int hb_hdb_split(const int32_t *group_len, int32_t n_groups, int32_t gpu_seqs, uint64_t bytes,
                 int32_t blocksx, int32_t threadsx, int32_t cpu_threads, float *ratio,
                 int32_t *fixed, hb_hdb_share *out);

/* The ratio for volume + 1 after volume's GPU results came back
 * (get_new_gpudb_ratio, functions.cpp:51-140): the slowest CPU thread's
 * time on the previous volume and the GPU's time on this one, each scaled
 * to the work they will do next, move the ratio by a share of their
 * relative difference.  Rings as the reference keeps them: volume_bytes[i
 * % 3] = .hDb bytes + residues left to the CPU of volume i; ratio_ring[i &
 * 1] = the ratio volume i ran with; hsp_ring[i & 1] = the GPU's HSPs on
 * volume i (sum of hb_subject_stats.n_hsp).  HB_ERR_ARG where the reference
 * exits (a zero volume size or ring ratio) and for volume 0. */
typedef struct hb_balance_in {
    uint32_t volume;            /* the volume whose GPU results just came back (>= 1) */
    float ratio;                /* the current ratio */
    uint32_t cpu_threads, gpu_cards;
    uint64_t volume_bytes[3];
    float ratio_ring[2];
    int64_t hsp_ring[2];
    int32_t gapped;             /* 1: a gapped search follows */
    int32_t query_length;       /* concatenated query length */
    uint32_t avg_query_length;  /* query length / contexts (functions.cpp:1280-1286) */
    double gpu_seconds;         /* device time of volume's scan */
    double cpu_seconds;         /* slowest CPU thread on volume - 1's CPU share */
} hb_balance_in;
int hb_balance_next_ratio(const hb_runtime_options *o, const hb_balance_in *in, float *ratio);

/* --- memory fit and the co-scheduled volume loop -------------------------
 * Device bytes a plain scan of q over a database (or a .hDb volume share)
 * needs, and whether they fit: replaces H_BLAST_check_memory
 * (H-BLAST/h_blast.cu:1280-1349, declared at gpu_cpu_common.h:93-102), which
 * sums the volume, lookup table, diagonal arrays and result buffers against
 * the device's memory.  Here: the upload (file_bytes of a .hDb share, then
 * its residues; file_bytes = 0 for packed arrays) with the per-sequence
 * metadata, the scan's copy of the residues, the HSP reservation, the
 * per-subject counters and the result layout, and the diagonal state if q
 * keeps it in global memory for subjects up to max_len.  residues = 0 takes
 * file_bytes as its bound.  budget < 0 takes the device's free memory.
 * *fits = 1 if *need <= the budget.  (The lookup table of q is already on
 * the device.) */
int hb_check_memory(hb_ctx *ctx, const hb_query *q, uint64_t file_bytes, uint64_t residues,
                    uint32_t n_seqs, uint64_t max_len, int64_t budget, uint64_t *need,
                    int32_t *fits);

/* One volume of the co-scheduled loop (GPU_BLAST_PreliminarySearchEngine,
 * ncbi_blast_files/functions.cpp:1409-1790). */
typedef struct hb_volume_report {
    int64_t oid_base;         /* database oid of the volume's first sequence */
    int32_t volume_seqs;      /* sequences in the volume */
    int32_t gpu_seqs;         /* the device's share (a prefix of the .hDb); 0 if none */
    float ratio;              /* the workload ratio the split used */
    int32_t on_cpu;           /* 1: the share did not fit the device, the CPU searched it all */
    uint64_t need_bytes;      /* device bytes the share needed (hb_check_memory) */
    double gpu_seconds;       /* device time of the share's scan */
    double cpu_seconds;       /* wall time of the CPU search of the rest */
    uint64_t n_hsp;           /* HSPs of the device share */
    uint64_t launch_seq;      /* order of the share's launch among the loop's launches and
                                 CPU searches (0: not launched) */
    uint64_t cpu_seq;         /* order of the volume's CPU search */
} hb_volume_report;

/* The caller's CPU side (NCBI's WordFinder threads, blast_engine.c:555-559,
 * over functions.cpp:1590-1768): search database oids [oid_lo, oid_hi) of
 * volume `volume`; nonzero stops the loop. */
typedef int (*hb_cpu_search_fn)(void *user, int32_t volume, int64_t oid_lo, int64_t oid_hi);
/* The device share's results of one volume, handed over before that
 * volume's CPU search: per-subject stats of its n_subjects sequences and its
 * HSPs, whose oid fields are database oids; nonzero stops the loop. */
typedef int (*hb_gpu_results_fn)(void *user, int32_t volume, int64_t oid_base,
                                 const hb_subject_stats *stats, uint32_t n_subjects,
                                 const hb_hsp *hsps, uint64_t n_hsp);
/* Device steps of the loop.  hb_cosearch binds the library's own (upload
 * with hb_db_upload_hdb, hb_scan_start, hb_scan_results); another binding
 * can drive the same loop (the tests drive it on the CPU).  fits: device
 * bytes the share needs and whether they fit; launch: upload the share (the
 * .hDb prefix in hdb) and start its scan, HB_ERR_NOMEM to send the volume to
 * the CPU; fetch: wait for the started share, hand its results on, return
 * its device seconds and HSP count; abandon: drop a started share after an
 * error. */
typedef struct hb_cosched_steps {
    void *user;
    int (*fits)(void *user, int32_t volume, const hb_hdb_share *share, int32_t max_len,
                uint64_t *need, int32_t *fits);
    int (*launch)(void *user, int32_t volume, const uint8_t *hdb, const hb_hdb_share *share,
                  int32_t blocksx, int32_t threadsx, const int32_t *group_len);
    int (*fetch)(void *user, int32_t volume, int64_t oid_base, double *gpu_seconds,
                 uint64_t *n_hsp);
    void (*abandon)(void *user);
} hb_cosched_steps;
/* Runs the loop over the volumes of the .hDbI at index_path (volume files
 * beside it): for v = 0, 1, ...: fetch volume v's device results, update the
 * ratio (hb_balance_next_ratio, from the second volume on unless fixed),
 * split volume v + 1 (hb_hdb_split), check it fits (else the CPU takes it
 * whole) and launch it, then run the CPU search of volume v's rest.
 * options NULL = the defaults; reports receives one entry per volume
 * (report_cap >= the volume count, else HB_ERR_CAPACITY with *n_volumes
 * set).  A callback's nonzero return stops the loop with HB_ERR_STATE. */
int hb_cosched_run(const char *index_path, const hb_runtime_options *options,
                   int32_t cpu_threads, int32_t gapped, int32_t query_length,
                   uint32_t avg_query_length, const hb_cosched_steps *steps,
                   hb_cpu_search_fn cpu, void *cpu_user, hb_volume_report *reports,
                   int32_t report_cap, int32_t *n_volumes);
/* The co-scheduled search of one query (as hb_query_prepare_masked takes it) over
 * the volumes of an .hDbI: the device scans each volume's share (uploaded
 * with hb_db_upload_hdb) while cpu(user, v, lo, hi) searches the rest of
 * volume v - 1, as H-BLAST pipelines them; gpu(user, ...) receives each
 * share's results.  device_budget: bytes a share may use on the device (< 0:
 * the free memory at each share); a share over it, or whose upload runs out
 * of memory, goes to the CPU whole. */
int hb_cosearch(hb_ctx *ctx, const char *index_path, const uint8_t *seq, int32_t len,
                int32_t n_ctx, const int32_t *ctx_off, const int32_t *ctx_len,
                const int32_t *xdrop, const int32_t *cutoff, const hb_params *params,
                const uint8_t *lut_mask, const hb_runtime_options *options,
                int32_t cpu_threads, int32_t gapped, int64_t device_budget,
                hb_cpu_search_fn cpu, hb_gpu_results_fn gpu, void *user,
                hb_volume_report *reports, int32_t report_cap, int32_t *n_volumes);

/* --- SEG low-complexity masking of a query ------------------------------
 * NCBI BLAST+ masks BLASTX queries (and BLASTP queries with -seg) with SEG
 * before the lookup table is built (-seg 12 2.2 2.5,
 * ncbi_blast_files/blast_args.cpp:309-316); NCBI's blast_seg.c is not part
 * of the reference, so hb_seg_mask restates the published algorithm
 * (Wootton and Federhen 1993, 1996) with NCBI's protein parameters.  Its
 * mask feeds hb_query_prepare_masked / hb_batch_add_masked (no word seeds
 * inside a segment); hard masking also replaces the residues by X (21). */
typedef struct hb_seg_params {
    int32_t window;     /* window length (12) */
    double locut;       /* trigger: window entropy <= locut bits (2.2) */
    double hicut;       /* extension: window entropy <= hicut bits (2.5) */
    int32_t maxbogus;   /* windows with more non-standard letters get no entropy (2) */
    int32_t maxtrim;    /* trimming tries at most this many shorter lengths (100) */
} hb_seg_params;
void hb_seg_params_default(hb_seg_params *p);
/* mask[i] = 1 for residues inside a low-complexity segment of their
 * context, else 0 (len bytes).  params NULL: the defaults.  HB_ERR_RESIDUE
 * for codes >= 28. */
int hb_seg_mask(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                const int32_t *ctx_len, const hb_seg_params *params, uint8_t *mask);

/* --- plain per-query scan ----------------------------------------------
 * hb_scan_start launches the stage on the context's stream and returns
 * (asynchronous, like H_BLAST_pre, ncbi_blast_files/gpu_cpu_common.h:48-54).
 * hb_scan_results blocks until the scan is done and copies results out
 * (like H_BLAST_get_data, gpu_cpu_common.h:56-60): stats gets
 * hb_db_subjects() entries, hsps gets every HSP ordered by
 * (oid, s_off, q_off) -- the order the CPU word finder discovers them --
 * and *n_hsp the total.  If cap is too small, *n_hsp is still set and
 * HB_ERR_CAPACITY is returned; call again with a larger array. */
int hb_scan_start(hb_ctx *ctx, hb_db *db, hb_query *q, hb_scan **out);
int hb_scan_wait(hb_scan *s);
int hb_scan_results(hb_scan *s, hb_subject_stats *stats, hb_hsp *hsps,
                    uint64_t cap, uint64_t *n_hsp);

/* The compact form of the same results, for a consumer that, like
 * blast_engine.c:540-586, needs per-subject data only for the subjects with
 * HSPs and the ungapped statistics only as sums (Blast_UngappedStatsUpdate
 * accumulates lookup hits and extensions).  The reference's get_data copies
 * 16 B per database subject (h_Hits, gpu_cpu_common.h:56-60); this copies
 * 24 B per subject with HSPs.  subj gets one entry per subject with HSPs in
 * oid order (first: index of its first record in hsps), hsps every HSP as
 * hb_scan_results orders them, totals the sums over all subjects.  There
 * are at most n_hsp subjects with HSPs (hb_scan_results with NULL arrays
 * gives n_hsp), so arrays of n_hsp entries always suffice; with smaller
 * capacities totals is still filled and HB_ERR_CAPACITY returned. */
typedef struct hb_subject_hits {
    uint32_t oid;
    uint32_t n_hsp;
    uint32_t lookup_hits;
    uint32_t extensions;
    uint64_t first;
} hb_subject_hits;
typedef struct hb_scan_totals {
    uint64_t lookup_hits;    /* word hits over all subjects */
    uint64_t extensions;     /* two-hit extensions over all subjects */
    uint64_t n_hsp;          /* HSPs */
    uint64_t n_subjects;     /* subjects with HSPs */
} hb_scan_totals;
int hb_scan_results_compact(hb_scan *s, hb_subject_hits *subj, uint64_t subj_cap,
                            hb_hsp *hsps, uint64_t hsp_cap, hb_scan_totals *totals);
/* Device time of the scan kernel(s) in milliseconds (hipEvent pair on the
 * stream the kernels run on); replaces get_gpu_walltime
 * (gpu_cpu_common.h:71). */
int hb_scan_elapsed_ms(hb_scan *s, float *ms);
int hb_scan_free(hb_scan *s);
/* Where the scan keeps its diagonal state: HB_STATE_LDS or HB_STATE_GLOBAL
 * (see hb_query_set_state_mode). */
int hb_scan_state_mode(const hb_scan *s);

/* No counterpart in the reference: a plain scan that allocates nothing, for
 * callers that cannot call hipMalloc (code under HIP graph capture, tight
 * serving loops, harnesses that forbid allocation).  It launches the same
 * kernels as hb_scan_start on `stream` (a hipStream_t; NULL: the context's
 * stream) and
 * returns at once:
 *   d_hsps    - device array of hsp_cap records: the scan's HSPs, unordered
 *               (sort by (oid, s_off, q_off) for hb_scan_results' order);
 *               records past hsp_cap are counted but not stored;
 *   d_counts  - device array of 3 * n_subjects uint32, per oid {lookup hits,
 *               extensions, HSPs}, or NULL;
 *   d_n_hsp   - device uint32: the number of HSPs found (may exceed hsp_cap:
 *               then run again with a larger array).
 * Its counters, and the diagonal state of a query that keeps it in global
 * memory (HB_NOALLOC_STATE_BYTES at most, with fewer workgroups for long
 * queries), live in a fixed scratch array in device memory, one per GPU,
 * held until the process ends; a query whose state does not fit one
 * workgroup's share returns HB_ERR_UNSUPPORTED.  Every call on a GPU shares
 * that GPU's scratch: calls that overlap on one GPU (on different streams or
 * threads) race and must not be issued. */
#define HB_NOALLOC_STATE_BYTES (4u << 20)
int hb_scan_noalloc(hb_ctx *ctx, hb_db *db, hb_query *q, hb_hsp *d_hsps, uint32_t hsp_cap,
                    uint32_t *d_counts, uint32_t *d_n_hsp, void *stream);

/* Convenience: prepare + start + results for one query. */
int hb_search(hb_ctx *ctx, hb_db *db, const uint8_t *seq, int32_t len,
              int32_t n_ctx, const int32_t *ctx_off, const int32_t *ctx_len,
              const int32_t *xdrop, const int32_t *cutoff, const hb_params *params,
              hb_subject_stats *stats, hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp);

/* --- preliminary gapped extension ---------------------------------------
 * The consumer of the seed stage's HSPs, run on the device for every subject
 * of a finished scan at once.  Replaces the per-subject calls
 * aux_struct->GetGappedScore(...) = BLAST_GetGappedScore
 * (ncbi_blast_files/blast_engine.c:617, blast_gapalign.c:3233-3560) made by
 * the CPU threads on the HSPs of H_BLAST_get_data, for in-frame protein
 * searches (BLASTP, BLASTX), score only, without restricted alignment:
 * per subject the ungapped HSPs are taken by descending score
 * (Blast_InitHitListSortByScore, blast_engine.c:584), an HSP held by a saved
 * gapped alignment of at least its score is skipped (the interval-tree test,
 * blast_gapalign.c:3402), the others are extended both ways from the middle
 * of their best 11-residue window (blast_gapalign.c:3184-3229) by the
 * X-drop DP of Blast_SemiGappedAlign (blast_gapalign.c:711-937), and saved
 * when they reach the context's gapped cutoff (blast_gapalign.c:3518-3534). */
typedef struct hb_gap_params {
    int32_t gap_open;    /* default 11 (BLAST_GAP_OPEN_PROT, blast_options.h:81) */
    int32_t gap_extend;  /* default 1 (BLAST_GAP_EXTN_PROT, blast_options.h:88) */
    int32_t x_dropoff;   /* raw gapped X-drop; default 38 */
    int32_t reserved;    /* 0 */
} hb_gap_params;
/* Defaults: 11, 1 and X = (int)(15 ln2 / 0.267) = 38, the 15-bit
 * BLAST_GAP_X_DROPOFF_PROT (blast_options.h:118) with BLOSUM62 11/1 gapped
 * lambda 0.267 (truncated, as BlastExtensionParametersNew converts it). */
void hb_gap_params_default(hb_gap_params *p);
#define HB_DEFAULT_GAP_XDROP 38

/* One gapped HSP (the fields of Blast_HSPInit, blast_gapalign.c:3522-3528). */
typedef struct hb_gapped_hsp {
    uint32_t oid;
    int32_t context;
    int32_t q_start, q_end;   /* context-relative query range, end exclusive */
    int32_t s_start, s_end;   /* subject range, end exclusive */
    int32_t q_gapped_start;   /* start point of the extension (context-relative) */
    int32_t s_gapped_start;
    int32_t score;            /* gapped raw score */
    int32_t pad;
} hb_gapped_hsp;

/* Per-subject counters of the gapped stage. */
typedef struct hb_gapped_stats {
    uint32_t extensions;  /* gapped extensions performed (gapped_stats->extensions) */
    uint32_t n_hsp;       /* gapped HSPs saved (score >= the context's cutoff) */
    int64_t first;        /* index of the subject's first gapped HSP, -1 if none */
} hb_gapped_stats;

/* Runs the gapped stage on the HSPs of scan s (waiting for the scan first).
 * gap_cutoff holds one raw gapped cutoff per query context
 * (hit_params->cutoffs[context].cutoff_score).  hsps receives the gapped
 * HSPs grouped by ascending oid, each subject's in save order; a capacity
 * equal to the scan's HSP count always suffices.  stats (hb_db_subjects()
 * entries) may be NULL. */
int hb_scan_gapped(hb_scan *s, const hb_gap_params *gp, const int32_t *gap_cutoff,
                   hb_gapped_stats *stats, hb_gapped_hsp *hsps, uint64_t cap,
                   uint64_t *n_hsp);
/* Device time of the last hb_scan_gapped's kernels (ms) and the number of
 * subjects it handed to the wide (global-memory) pass. */
int hb_scan_gapped_info(hb_scan *s, float *ms, uint32_t *n_wide);
/* The same stage on ungapped HSPs the caller holds (any order): those of
 * hb_batch_results / hb_serve_results for one query, or NCBI's own initial
 * hit lists.  The query is given as to hb_query_prepare (matrix NULL =
 * BLOSUM62); in[i].context indexes its contexts and q_start / q_off are
 * concatenated-query offsets.  Each record must lie inside its context and
 * subject with q_off - s_off == q_start - s_start (else HB_ERR_ARG).
 * Outputs as hb_scan_gapped; a capacity of n_in always suffices. */
int hb_gapped_search(hb_ctx *ctx, hb_db *db, const uint8_t *seq, int32_t len, int32_t n_ctx,
                     const int32_t *ctx_off, const int32_t *ctx_len, const int8_t *matrix,
                     const hb_hsp *in, uint64_t n_in, const hb_gap_params *gp,
                     const int32_t *gap_cutoff, hb_gapped_stats *stats, hb_gapped_hsp *hsps,
                     uint64_t cap, uint64_t *n_hsp);

/* --- batched form --------------------------------------------------------
 * Many queries (BLASTP or six-frame BLASTX, mixed lengths) searched against
 * the resident database in one persistent launch per size class.  Queries
 * are packed into groups of about group_residues concatenated residues,
 * separated by `window` code-0 residues so that hits of different queries
 * never pair (each query's HSPs equal its HSPs searched alone); each group
 * has its own lookup table in one device arena, and workgroups pop (group,
 * tile chunk) items from one queue.  The launch sequence is captured in a
 * hipGraph on the first run and replayed on later runs.  H-BLAST's closest
 * counterpart is NCBI's concatenation of a query batch into one query
 * (contexts) searched by H_BLAST_pre/H_BLAST_get_data
 * (h_blast.cu:922-1157); here the concatenation is split into groups that
 * fit on-chip memory. */
typedef struct hb_batch hb_batch;
typedef struct hb_batch_opts {
    int32_t group_residues;  /* target concatenated residues per group; 0 = autotune this
                                and the launch geometry below */
    int32_t use_graph;       /* 1: capture the launches in a hipGraph and replay it */
    uint32_t tune_tiles;     /* database tiles timed per autotune candidate (0 = default) */
    /* Launch geometry (the reference exposes its grid as -gpu_blocks /
     * -gpu_threads, cmdline_flags.cpp:213-230): 0 = autotuned with the group
     * size, or the default when group_residues is given. */
    int32_t waves_per_workgroup;   /* 4, 2 or 1 (default: the most wavefronts per CU) */
    int32_t ring_bytes;            /* staging ring per wavefront (multiple of 64, >= 256;
                                      default: the largest that keeps the occupancy) */
    int32_t items_per_workgroup;   /* (group, tile chunk) work items per workgroup, which
                                      sets the subjects per item (default 64) */
} hb_batch_opts;

typedef struct hb_batch_info {
    int32_t n_queries;
    int32_t n_groups;
    int32_t n_classes;       /* launches (groups of one LDS configuration) */
    int32_t group_residues;  /* chosen (autotuned or given) */
    float kernel_ms;         /* last run: hipEvent time of the launches */
    float build_ms;          /* grouping + lookup tables (host clock) */
    uint64_t lookup_hits;    /* last run, all queries */
    uint64_t extensions;
    uint64_t hsps;
    int32_t waves_per_workgroup;   /* chosen launch geometry (of the largest class) */
    int32_t ring_bytes;
    int32_t items_per_workgroup;
    int32_t tune_candidates;       /* autotune candidates timed (0: none) */
} hb_batch_info;

void hb_batch_opts_default(hb_batch_opts *o);
int hb_batch_create(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_batch_opts *opts,
                    hb_batch **out);
/* Adds one query (same arguments as hb_query_prepare); *query_id is its
 * index in the batch.  Queries can be added until the first hb_batch_run. */
int hb_batch_add(hb_batch *b, const uint8_t *seq, int32_t len, int32_t n_ctx,
                 const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                 const int32_t *cutoff, int32_t *query_id);
/* hb_batch_add with a lookup-table mask (as hb_query_prepare_masked). */
int hb_batch_add_masked(hb_batch *b, const uint8_t *seq, int32_t len, int32_t n_ctx,
                        const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                        const int32_t *cutoff, const uint8_t *lut_mask, int32_t *query_id);
/* Builds groups and lookup tables on the first call (autotuning the group
 * size if asked), then runs the scan; blocking. */
int hb_batch_run(hb_batch *b);
/* HSPs of one query from the last run, in (oid, s_off, q_off) order, with
 * context and query offsets relative to that query alone. */
int hb_batch_results(hb_batch *b, int32_t query_id, hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp);
int hb_batch_info_get(hb_batch *b, hb_batch_info *info);
int hb_batch_free(hb_batch *b);

/* --- served form ---------------------------------------------------------
 * A request loop over the resident database: queries are submitted one at
 * a time from any thread, admitted by a device-memory budget, and folded
 * into batched scans (the batched form above, fixed group size) that run on
 * their own streams, up to max_inflight at once, so that a batch formed
 * while another is on the GPU starts as soon as device resources free up.
 * Each query's HSPs are identical to those of the query searched alone.
 * No counterpart in the reference: an H-BLAST process searches the query
 * batch it was started with (H_BLAST_Execute, functions.cpp:1167-1217). */
typedef struct hb_server hb_server;
typedef struct hb_serve_opts {
    int32_t max_batch_queries;   /* queries per batch (default 64) */
    int32_t max_inflight;        /* batches on the device at once (default 2) */
    int64_t max_batch_residues;  /* query residues per batch (default 16384) */
    int64_t mem_budget;          /* device bytes for admitted batches; 0 = half the free memory */
    int32_t group_residues;      /* group size of the batched scans (default 768) */
    int32_t max_pending;         /* queued requests beyond which submit is refused (default 65536) */
    double hsp_per_cell;         /* expected HSPs per (query residue x subject residue), for
                                    the output reservation (default 5e-7) */
    int64_t linger_us;           /* a batch that is not full waits this long after its oldest
                                    request for more (default 0: launch at once) */
} hb_serve_opts;

typedef struct hb_serve_stats {
    uint64_t submitted, completed, rejected, batches;
    uint64_t query_residues;     /* completed */
    uint64_t peak_bytes;         /* largest admitted device footprint */
    int32_t peak_inflight;       /* most batches on the device at once */
    int32_t reruns;              /* batches rerun because their output overflowed */
    double p50_ms, p99_ms;       /* submit -> results ready, over the most recent 65,536
                                    completed queries */
    double mean_ms, max_ms;      /* ... over all completed queries */
    double gpu_ms;               /* sum of batch kernel times */
    double wall_ms;              /* since start */
} hb_serve_stats;

void hb_serve_opts_default(hb_serve_opts *o);
int hb_serve_start(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_serve_opts *opts,
                   hb_server **out);
/* Queues one query (same arguments as hb_query_prepare).  Returns
 * HB_ERR_BUSY when max_pending requests are queued and HB_ERR_CAPACITY when
 * the query alone exceeds the memory budget; *ticket names the request. */
int hb_serve_submit(hb_server *s, const uint8_t *seq, int32_t len, int32_t n_ctx,
                    const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                    const int32_t *cutoff, uint64_t *ticket);
/* As hb_serve_submit with a lookup-table mask (as hb_query_prepare_masked):
 * a served BLASTX request masked by SEG, as BLAST+ masks it by default. */
int hb_serve_submit_masked(hb_server *s, const uint8_t *seq, int32_t len, int32_t n_ctx,
                           const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                           const int32_t *cutoff, const uint8_t *lut_mask, uint64_t *ticket);
/* Waits up to timeout_ms (< 0: no limit) for the request: HB_OK when its
 * results are ready, HB_ERR_BUSY on time-out, the batch's error otherwise. */
int hb_serve_wait(hb_server *s, uint64_t ticket, int32_t timeout_ms);
/* Results of a finished request in (oid, s_off, q_off) order.  With hsps
 * NULL only *n_hsp is set; a successful copy releases the ticket, and so
 * does returning the error of a request that failed. */
int hb_serve_results(hb_server *s, uint64_t ticket, hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp);
int hb_serve_stats_get(hb_server *s, hb_serve_stats *st);
/* Finishes every queued and running request, then stops the loop. */
int hb_serve_stop(hb_server *s);

/* --- diagnostics ---------------------------------------------------------
 * Copies `bytes` between two fresh device buffers `reps` times with a plain
 * 16-byte-per-lane copy kernel and returns the mean time per copy: the
 * known-size kernel used to calibrate rocprof byte counters and to measure
 * the achievable HBM bandwidth.  No counterpart in the reference (its
 * timing is get_gpu_walltime, gpu_cpu_common.h:71). */
int hb_diag_copy(hb_ctx *ctx, uint64_t bytes, int reps, float *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* HB_SEED_H */
