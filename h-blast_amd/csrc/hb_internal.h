// hb_internal.h -- structures shared by the host runtime and the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hb_seed.h"

#define HB_ALPHA 28
#define HB_NCELLS 32768     // 3 letters x 5 bits
#define HB_PV_WORDS 1024    // HB_NCELLS / 32
#define HB_WAVE 64
#define HB_JOBCAP 128       // queued two-hit extensions per wave
#define HB_BLKCAP 128       // hits waiting on a pending diagonal per wave
#define HB_SCAP 1024        // subjects up to this length are staged in LDS
#define HB_BLOCK 256        // 4 wavefronts per workgroup

// Kernel arguments of the scan (one query group against one database).
struct hb_scan_args {
    // database (resident)
    const uint8_t *res;          // all residues
    const uint64_t *offs;        // n+1 subject offsets
    const uint32_t *order;       // processing order (length-descending oids)
    const uint32_t *tiles;       // n_tiles+1 ranges into order
    uint32_t n_tiles;
    // query group
    const uint2 *cells;          // HB_NCELLS cells {start, count}
    const uint32_t *entries;     // query offsets
    uint32_t n_entries;
    const uint32_t *pv;          // HB_PV_WORDS presence bits
    const uint8_t *query;        // qlen_pad bytes (zero padded)
    int32_t qlen;
    int32_t qlen_pad;
    const int32_t *ctx_off;      // n_ctx
    const int32_t *ctx_xdrop;
    const int32_t *ctx_cutoff;
    int32_t n_ctx;
    int32_t ctx_pad;             // n_ctx rounded up to 4 (LDS table stride)
    const int8_t *mat32;         // 32x32 scores, -128 padded
    int32_t window;
    int32_t diag_len;            // >= qlen + window, multiple of 64 (diagonal = (q - s) mod diag_len)
    uint32_t wave_lds;           // LDS bytes per wave
    // work queue and output
    uint32_t *tile_counter;
    hb_hsp *out;
    uint32_t out_cap;
    uint32_t *out_count;
    uint32_t *stats;             // 3 per oid: lookup hits, extensions, HSPs
};

// Kernel launchers (hb_kernels.hip).
hipError_t hb_launch_check_residues(const uint8_t *res, uint64_t n, uint32_t *flag,
                                    hipStream_t st);
hipError_t hb_launch_lut_count(const uint8_t *query, int32_t qlen, const int32_t *word_ok,
                               const int8_t *mat32, int32_t threshold, uint32_t *count,
                               hipStream_t st);
hipError_t hb_launch_lut_scan(uint32_t *count_to_start, uint32_t *total, hipStream_t st);
hipError_t hb_launch_lut_fill(const uint8_t *query, int32_t qlen, const int32_t *word_ok,
                              const int8_t *mat32, int32_t threshold, const uint32_t *start,
                              uint32_t *entries, uint2 *cells, uint32_t *pv, hipStream_t st);
hipError_t hb_launch_scan(const hb_scan_args &a, int n_blocks, int block_threads,
                          size_t lds_bytes, hipStream_t st);
hipError_t hb_scan_occupancy(size_t lds_bytes, int block_threads, int *blocks_per_cu);
hipError_t hb_launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t st);
