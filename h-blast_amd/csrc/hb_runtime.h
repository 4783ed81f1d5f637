// hb_runtime.h -- host-side objects shared by the plain (hb_seed.cpp) and
// batched/served (hb_batch.cpp, hb_serve.cpp) entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "hb_internal.h"

void hb_note_hip_error(hipError_t e, const char *what, int line);

#define HB_CHECK(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) {                          \
            hb_note_hip_error(e_, #call, __LINE__);      \
            return HB_ERR_DEVICE;                        \
        }                                                \
    } while (0)

extern const int8_t kHbBlosum62[HB_ALPHA][HB_ALPHA];

struct hb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int n_cu = 0;
    size_t max_lds = 0;
    std::mutex mu;
    // Device buffers of finished scans, reused by later scans so that a
    // per-query loop does not pay hipMalloc/hipFree per call.
    std::mutex pool_mu;
    std::vector<std::pair<void *, size_t>> pool;
    size_t pool_bytes = 0;
    // pinned host staging for result copies (grown on demand, kept)
    std::mutex pin_mu;
    void *pinned = nullptr;
    size_t pinned_bytes = 0;
    // device address of this GPU's hb_scan_noalloc scratch
    unsigned char *scratch_dev = nullptr;
};

// Copies device bytes to pageable host memory through the context's pinned
// staging buffer (DMA at full link speed, then a multi-threaded host copy).
int hb_copy_out(hb_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t st);

void *hb_pool_get(hb_ctx *c, size_t bytes);
void hb_pool_put(hb_ctx *c, void *p, size_t bytes);

struct hb_db {
    hb_ctx *ctx = nullptr;
    uint32_t n = 0;
    uint64_t total = 0;
    uint8_t *d_res = nullptr;
    uint64_t *d_offs = nullptr;
    uint4 *d_subj = nullptr;     // {oid, length, offset into d_scan_res} in processing order
    uint8_t *d_scan_res = nullptr;  // residues in processing order: the scan's copy (each
                                    // subject 16-byte aligned, spans padded to 16 bytes)
    uint64_t scan_bytes = 0;        // bytes of that copy's subject spans
    uint32_t *d_tiles = nullptr;
    uint32_t n_tiles = 0;
    std::vector<uint64_t> h_offs;
    uint64_t max_len = 0;   // longest subject
    float hdb_len_ms = 0.f, hdb_copy_ms = 0.f;   // hb_db_upload_hdb's kernels
};

// Search parameters resolved from hb_params.
struct hb_resolved_params {
    int32_t threshold = 11;
    int32_t window = 40;
    int8_t mat32[32 * 32];
};
int hb_resolve_params(const hb_params *p, hb_resolved_params *out);

// A concatenated query as built on the host: residues and per-context
// (offset, length, X-drop, cutoff).
struct hb_concat {
    std::vector<uint8_t> seq;
    std::vector<int32_t> ctx_off, ctx_len, xdrop, cutoff;
    std::vector<uint8_t> mask;   // lookup-table mask per residue (empty: none)
};
// Validates contexts and residues of one query.
int hb_check_query(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                   const int32_t *ctx_len, const int32_t *xdrop);

// Device lookup tables and group descriptors for a set of concatenated
// queries, all in two allocations (fixed-size tables + entries), taken from
// and returned to the context's buffer pool.
struct hb_group_set {
    hb_ctx *ctx = nullptr;
    std::vector<hb_group_desc> desc;   // host copy
    hb_group_desc *d_desc = nullptr;   // device copy
    void *d_fixed = nullptr;           // cells, starts, pv, query, ctx, qflag per group
    uint32_t *d_entries = nullptr;
    int8_t *d_mat = nullptr;
    size_t fixed_bytes = 0, entries_bytes = 0;
    int32_t qlen_pad_max = 0, ctx_pad_max = 4, diag_len_max = 64;
    std::vector<uint32_t> n_entries, n_cells_used;
};
int hb_group_set_build(hb_ctx *c, const std::vector<hb_concat> &queries,
                       const std::vector<int32_t> &ctx_base, const hb_resolved_params &p,
                       hipStream_t st, hb_group_set *out);
void hb_group_set_free(hb_group_set *g);

// Launch configuration of the scan for a given LDS layout.
struct hb_launch_cfg {
    int waves = 4;
    size_t lds = 0;
    int blocks = 0;
    int32_t ring = 1024;   // staging ring bytes per wave
};
// Where a launch keeps its per-diagonal two-hit state: 32-bit words in LDS,
// or 32-bit words in a global-memory region per wavefront (markers in LDS).
enum hb_state_place { HB_PLACE_LDS = 0, HB_PLACE_GLOBAL = 2 };
static inline int hb_state_bytes(int place) { return place == HB_PLACE_LDS ? 4 : 0; }
// LDS words hold 23-bit positions between resets, so a subject and its
// successor must fit: 2 * (longest subject + window) + 64 < 2^23.
static inline bool hb_state_lds_ok(uint64_t max_len, int32_t window)
{
    return 2 * (max_len + (uint64_t)window) + 64 < (uint64_t)HB_LDS_POS_LIMIT;
}
// The reset threshold of the state (hb_scan_args::state_limit).
static inline int32_t hb_state_limit(int place, uint64_t max_len)
{
    return place == HB_PLACE_LDS ? HB_LDS_POS_LIMIT - (int32_t)max_len : HB_GLOBAL_POS_LIMIT;
}
// Automatic placement for diagonal arrays of D entries and `fixed` bytes of
// per-workgroup LDS on a database whose longest subject is max_len: LDS when
// the subjects allow it and it leaves enough wavefronts per CU, else global
// memory.
int hb_state_choose(const hb_ctx *c, int32_t D, size_t fixed, uint64_t max_len, int32_t window);
// The scan keeps a concatenated query, its context tables and the score
// matrix in LDS (with the diagonal state in global memory at the least), and
// the lookup-table build keeps the query in LDS too: queries past about
// 150,000 concatenated residues do not fit a workgroup's 160 KB.
bool hb_query_lds_ok(const hb_ctx *c, int32_t qlen, int32_t n_ctx);
// Global-memory diagonal state of one launch is capped at this size: a
// launch for a very long query then runs fewer workgroups.
constexpr size_t HB_GSTATE_MAX_BYTES = size_t(1) << 30;
// Bytes per diagonal of the global-memory state.
constexpr size_t HB_GSTATE_WORD = 4;
// Bytes of that state for diagonal arrays of D entries (upper bound).
size_t hb_gstate_bytes_max(const hb_ctx *c, int32_t D);
// place: where the diagonal state goes (LDS bytes per diagonal follow).
// ring_cap > 0 caps the staging ring, waves > 0 fixes the wavefronts per
// workgroup (autotuner constraints; HB_ERR_UNSUPPORTED if nothing fits).
int hb_scan_config(hb_ctx *c, bool multi, int32_t qlen_pad_max, int32_t ctx_pad_max,
                   int32_t diag_len_max, uint32_t work_items, hb_launch_cfg *cfg,
                   int place = HB_PLACE_LDS, int32_t ring_cap = 0, int32_t waves = 0);

// Asynchronous steps of a batch (hb_batch.cpp), used by the serving loop:
// the batch runs on the given stream (nullptr: the context's); launch
// returns at once, ready polls, finish collects (rerunning the pass if the
// HSP buffer was too small).
struct hb_batch;
int hb_batch_create_on(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_batch_opts *opts,
                       hipStream_t st, hb_batch **out);
bool hb_batch_rerun_needed(const hb_batch *b);
int hb_batch_launch(hb_batch *b);
bool hb_batch_ready(hb_batch *b);
int hb_batch_finish(hb_batch *b);
size_t hb_batch_device_bytes(const hb_batch *b);
int hb_batch_reserve(hb_batch *b, uint32_t cap);
