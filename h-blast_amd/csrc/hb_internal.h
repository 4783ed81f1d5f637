// hb_internal.h -- structures shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hb_seed.h"

#define HB_ALPHA 28
#define HB_NCELLS 32768     // 3 letters x 5 bits
#define HB_PV_WORDS 1024    // HB_NCELLS / 32
#define HB_WAVE 64
#define HB_JOBCAP 128       // queued two-hit extensions per wave (a ring)
#define HB_BLKCAP 128       // hits waiting on a pending diagonal per wave
#define HB_SLOTS 8          // staged subjects per wave whose extensions may still be queued
#define HB_GMARKS 1024      // ordering-marker slots per wave when the diagonal words are in global memory
#ifndef HB_RING_MAX
#define HB_RING_MAX 4096    // largest staging ring per wave (subjects up to the ring are staged)
#endif
#define HB_BLOCK 256        // at most 4 wavefronts per workgroup
// Diagonal state kept in LDS: one 32-bit word per diagonal, 23-bit positions
// between resets (hb_scan_args::state_limit), so the longest subject and
// its successor must fit: 2 * (longest + window) + 64 < 2^23.
#define HB_LDS_POS_LIMIT (1 << 23)
// ... in global memory: one 32-bit word per diagonal (the ordering markers
// are in LDS), 29-bit positions (subjects up to 2^28 residues).
#define HB_GLOBAL_POS_LIMIT (1 << 28)

// One query group: a concatenated query (one or more queries / frames as
// contexts) with its lookup table.  The plain path scans one group; the
// batched path scans many groups in one launch.
struct hb_group_desc {
    const uint2 *cells;          // HB_NCELLS cells {start, count}
    const uint32_t *entries;     // query offsets, ascending per cell
    const uint32_t *pv;          // HB_PV_WORDS presence bits
    const uint8_t *query;        // qlen_pad bytes (zero padded)
    const int32_t *ctx;          // ctx_off[n_ctx] | xdrop[n_ctx] | cutoff[n_ctx]
    int32_t qlen;
    int32_t qlen_pad;            // multiple of 16
    int32_t n_ctx;
    int32_t diag_len;            // >= qlen + window, multiple of 64
    uint32_t n_entries;
    int32_t ctx_base;            // added to context indices in HSP records
    int32_t pad[2];
};

// Kernel arguments of the scan.
struct hb_scan_args {
    // database (resident)
    // the scan's copy of the residues, in processing order: each subject starts
    // 16-byte aligned and its span is padded to a multiple of 16 bytes
    const uint8_t *res;
    const uint64_t *offs;        // n+1 subject offsets
    // per subject in processing order (length-descending): {oid, length,
    // residue offset low, high}, read sequentially by the scan
    const uint4 *subj;
    const uint32_t *tiles;       // tile boundaries into subj
    uint32_t tile_begin;         // tiles [tile_begin, tile_end) are scanned
    uint32_t tile_end;
    uint32_t tiles_per_item;     // batched: tiles per (group, tile chunk) work item
    // query groups
    const hb_group_desc *groups;
    uint32_t n_groups;
    // LDS layout (maxima over the groups of the launch)
    int32_t qlen_pad_max;
    int32_t ctx_pad_max;         // multiple of 4
    int32_t diag_len_max;
    uint32_t wave_lds;           // LDS bytes per wave
    uint32_t ring;               // staging ring bytes per wave (multiple of 16): subjects up
                                 // to this length are staged in LDS
    const int8_t *mat32;         // 32x32 scores, -128 padded
    int32_t window;
    // work queue and output
    uint32_t *work_counter;
    hb_hsp *out;
    uint32_t out_cap;
    uint32_t *out_count;
    uint32_t *stats;             // plain: 3 per oid (lookup hits, extensions, HSPs); else null
    unsigned long long *group_totals;  // batched: 3 per group; else null
    // the diagonal state in global memory, diag_len_max 32-bit words per
    // wavefront of the grid (null: in LDS)
    uint32_t *gstate;
    // the state is reset when a subject's base offset + length + window
    // reaches this (the position limit of the placement - the longest subject)
    int32_t state_limit;
};

// per wave: diagonal words (or, with the words in global memory, the
// ordering-marker table), 64 segment tags, job and blocked queues (8 B
// entries), the staged subjects' slots (32 B each) and the staging ring.
// The regions of a workgroup's wavefronts are padded so that their distance
// modulo 512 bytes is at least 128: distances of 0 or 64 bytes (mod 512)
// made the C1 scan 8% slower on gfx950 (measured), which points to LDS bank
// or port conflicts between wavefronts.
// state_bytes: bytes per diagonal of LDS state (4, or 0 when the state is
// in global memory).
static inline size_t hb_scan_wave_lds(int32_t diag_len, int32_t ring, int state_bytes = 4)
{
    size_t b = (state_bytes ? (size_t)state_bytes * (size_t)diag_len : 4 * (size_t)HB_GMARKS) +
               256 + 8 * (size_t)(HB_JOBCAP + HB_BLKCAP) + 32 * HB_SLOTS + (size_t)ring;
    b = (b + 15) & ~(size_t)15;
    const size_t r = b % 512;
    if (r < 128) b += 128 - r;
    return b;
}
static inline int32_t hb_ctx_pad(int32_t n_ctx) { return (n_ctx + 3) & ~3; }
// per workgroup: score matrix, context tables, query
static inline size_t hb_scan_fixed_lds(int32_t qlen_pad, int32_t ctx_pad)
{
    return 1024 + 12 * (size_t)ctx_pad + qlen_pad;
}

// Kernel launchers (hb_kernels.hip).
// Copies every subject's residues (res at offs[oid]) to its processing-order
// place in out (the record's offset), one workgroup per tile.
hipError_t hb_launch_gather_subjects(const uint8_t *res, const uint64_t *offs, const uint4 *subj,
                                     const uint32_t *tiles, uint32_t n_tiles, uint8_t *out,
                                     hipStream_t st);
hipError_t hb_launch_check_residues(const uint8_t *res, uint64_t n, uint32_t *flag,
                                    hipStream_t st);
// Lookup tables of n groups (grid.y = group).  qflag[q] = residue | 0x80 when
// the 3-mer at q lies inside one context.
struct hb_lut_job {
    const uint8_t *qflag;
    int32_t qlen;
    int32_t pad;
    uint32_t *count_start;       // HB_NCELLS + 1: counts, then exclusive starts
    uint32_t *entries;
    uint2 *cells;
    uint32_t *pv;
    uint32_t *total;             // total entries of the group
};
// pass 0: count; pass 1: exclusive scan + total; pass 2: fill cells, entries, pv
hipError_t hb_launch_lut(const hb_lut_job *jobs_dev, int n_jobs, int32_t qlen_max,
                         const int8_t *mat32, int32_t threshold, int pass, hipStream_t st);
hipError_t hb_launch_scan(const hb_scan_args &a, bool multi, int n_blocks, int block_threads,
                          size_t lds_bytes, hipStream_t st);
// Sets the kernel's dynamic-LDS limit and returns its occupancy (call before
// launching or capturing a launch with that LDS size).
// Raises both scan kernels' dynamic LDS limit to max_lds (once per context).
hipError_t hb_scan_allow_lds(size_t max_lds);
hipError_t hb_scan_occupancy(bool multi, bool gstate, size_t lds_bytes,
                             int block_threads, int *blocks_per_cu);
hipError_t hb_launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t st);
hipError_t hb_launch_fill_u32(uint32_t *dst, uint32_t value, uint32_t n, hipStream_t st);

// Plain-form result layout on the device (hb_kernels.hip): per-subject stats
// with first-record indices, records grouped by subject and ordered by
// (s_off, q_off).  Subjects with more than HB_FIN_SORT_MAX records are left
// unordered and listed in big[] (count in *n_big) for the host.
#define HB_FIN_SORT_MAX 256
struct hb_fin_big {
    uint32_t oid, n;
    uint64_t first;
};
// scratch: ns + ns / 4096 + 1 uint32
hipError_t hb_launch_finalize(const uint32_t *st3, uint32_t ns, const hb_hsp *recs, uint32_t n,
                              hb_subject_stats *stats_out, hb_hsp *recs_out, uint32_t *scratch,
                              hb_fin_big *big, uint32_t big_cap, uint32_t *n_big, hipStream_t st);
// The compact layout: entries of the subjects with HSPs in oid order
// (subj_out, at most min(ns, n)), the records grouped and ordered as above,
// tot[0..1] += the lookup hits and extensions of all subjects (tot cleared
// by the caller).  With nb = (ns + 4095) / 4096, scratch holds 2 nb + 1 + ns
// uint32, and the number of subjects with HSPs ends up in scratch[2 nb].
hipError_t hb_launch_finalize_compact(const uint32_t *st3, uint32_t ns, const hb_hsp *recs,
                                      uint32_t n, hb_subject_hits *subj_out, hb_hsp *recs_out,
                                      uint32_t *scratch, unsigned long long *tot, hb_fin_big *big,
                                      uint32_t big_cap, uint32_t *n_big, hipStream_t st);
// First-record indices of per-subject counts st3[3i + 2] (hb_kernels.hip):
// first[i] = sum of the counts of subjects < i.  scratch: ns / 4096 + 1.
hipError_t hb_launch_first_index(const uint32_t *st3, uint32_t ns, uint32_t *first,
                                 uint32_t *scratch, hipStream_t st);

// ---------------------------------------------------------------- gapped stage
// Kernel arguments of the preliminary gapped extension (hb_gapped.hip).
#define HB_GAP_RING 256       // DP columns per wavefront kept in LDS (power of two)
#define HB_GAP_ORDER 256      // ungapped HSPs per subject ordered in LDS
#define HB_GAP_SAVED 32       // saved gapped HSPs per subject kept in LDS
#define HB_GAP_SUBJ 2048      // subjects up to this length are staged in LDS
struct hb_gap_args {
    const uint8_t *res;            // database residues
    const uint64_t *offs;          // n + 1 subject offsets
    const uint8_t *query;          // concatenated query
    const int32_t *ctx;            // ctx_off[n_ctx] | ctx_len[n_ctx] | cutoff[n_ctx]
    int32_t n_ctx;
    const int8_t *mat32;           // 32x32 scores
    int32_t gap_open, gap_extend, x_dropoff;
    const hb_subject_stats *ust;   // ungapped per-subject stats (first, n_hsp)
    const hb_hsp *urec;            // ungapped HSPs grouped by subject
    uint32_t n_subjects;
    uint32_t *counter;             // work counter (zeroed before the launch)
    const uint32_t *list;          // subjects with HSPs, most HSPs first
    const uint32_t *n_list;        // their number (device)
    hb_gapped_hsp *out;            // subject k's gapped HSPs at ust[k].first + i
    uint32_t *st3;                 // per subject {0, extensions, saved}
    uint32_t *wide;                // [0] count, then subjects for the wide pass
    // wide pass: subjects wide[1 .. n_wide], one per wavefront-iteration,
    // each wavefront with its own global scratch of scratch_words int32:
    // DP ring 2 * ring_wide, saved list 6 * n_max, order n_max
    uint32_t n_wide;
    int32_t *scratch;
    uint32_t ring_wide;            // power of two > longest subject + 1
    uint32_t n_max;                // most ungapped HSPs of a wide subject
    uint32_t scratch_words;
};
// Work list of the main pass: list[] = subjects with HSPs ordered by
// descending HSP count (capped at 63), *n_list their number; subjects without
// HSPs get zero stats.  scratch: 128 uint32 (zeroed here).
hipError_t hb_launch_gap_list(const hb_gap_args &a, uint32_t *list, uint32_t *n_list,
                              uint32_t *scratch, hipStream_t st);
hipError_t hb_launch_gapped(const hb_gap_args &a, bool wide, int n_blocks, hipStream_t st);
// st3[3 * oid + 2] += 1 for each of the n records (st3 zeroed by the caller).
hipError_t hb_launch_count_oids(const hb_hsp *rec, uint32_t n, uint32_t *st3, hipStream_t st);
// Dense layout: stats[k] and out_dense from the sparse output and first[].
hipError_t hb_launch_gap_compact(const hb_gap_args &a, const uint32_t *first,
                                 hb_gapped_stats *stats, hb_gapped_hsp *out_dense,
                                 hipStream_t st);

// Unpacking of a .hDb volume on the device (hb_hdb.hip).  Sequence c sits
// in block B = c / threadsx (global over the volume), group g = B / blocksx;
// its 4-byte stripe i is word (c % threadsx) + i * threadsx of the block,
// which starts at group_base[g] + (B % blocksx) * threadsx * group_len[g].
struct hb_hdb_args {
    const uint8_t *hdb;            // the volume's bytes
    const uint64_t *group_base;    // byte offset of each group
    const int32_t *group_len;      // each group's padded length (multiple of 16)
    uint32_t n_seqs, blocksx, threadsx;
    uint32_t *lens;                // out: sequence lengths (last non-zero byte + 1)
    const uint64_t *offs;          // packed offsets (copy pass)
    uint8_t *res;                  // out: packed residues (copy pass)
    uint32_t *bad;                 // out: set to 1 if a residue code is >= 28 (copy pass)
};
hipError_t hb_launch_hdb_lengths(const hb_hdb_args &a, hipStream_t st);
hipError_t hb_launch_hdb_copy(const hb_hdb_args &a, hipStream_t st);
