// hb_seed.cpp -- host runtime behind include/hb_seed.h.
//
// Replaces the host half of H-BLAST (h_blast_v1.1/H-BLAST/h_blast.cu):
// device setup, the resident database (instead of an mmap + H2D copy per
// call, h_blast.cu:587-653,1094), lookup-table staging (built on the device
// here instead of copied from NCBI's BlastAaLookupTable, h_blast.cu:940-1008),
// kernel launch (h_blast.cu:1109-1145), event timing (h_blast.cu:1101,1147)
// and result fetch/compaction (H_BLAST_get_data + memory_defrag,
// h_blast.cu:1160-1252).  Errors are returned, never exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

#include "hb_internal.h"

#define HB_CHECK(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) {                          \
            hb_note_hip_error(e_, #call, __LINE__);      \
            return HB_ERR_DEVICE;                        \
        }                                                \
    } while (0)

static thread_local char g_last_hip[256];

static void hb_note_hip_error(hipError_t e, const char *what, int line)
{
    snprintf(g_last_hip, sizeof(g_last_hip), "%s (%s) at hb_seed.cpp:%d", hipGetErrorString(e),
             what, line);
}

// BLOSUM62 in ncbistdaa order -,A,B,C,D,E,F,G,H,I,K,L,M,N,P,Q,R,S,T,V,W,X,Y,Z,U,*,O,J.
// Same values as the reference's table (h_blast_v1.1/H-BLAST/h_blast.c:42-71);
// the gap letter scores -128 so that context sentinels stop every extension.
static const int8_t kBlosum62[HB_ALPHA][HB_ALPHA] = {
#include "hb_blosum62.inc"
};

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
};

static void *pool_get(hb_ctx *c, size_t bytes)
{
    {
        std::lock_guard<std::mutex> lk(c->pool_mu);
        size_t best = (size_t)-1;
        for (size_t i = 0; i < c->pool.size(); ++i)
            if (c->pool[i].second >= bytes && c->pool[i].second <= 2 * bytes + 4096 &&
                (best == (size_t)-1 || c->pool[i].second < c->pool[best].second))
                best = i;
        if (best != (size_t)-1) {
            void *p = c->pool[best].first;
            c->pool.erase(c->pool.begin() + (long)best);
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}

static void pool_put(hb_ctx *c, void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(c->pool_mu);
    c->pool.emplace_back(p, bytes);
    if (c->pool.size() > 32) {
        hipFree(c->pool.front().first);
        c->pool.erase(c->pool.begin());
    }
}

struct hb_db {
    hb_ctx *ctx = nullptr;
    uint32_t n = 0;
    uint64_t total = 0;
    uint8_t *d_res = nullptr;
    uint64_t *d_offs = nullptr;
    uint32_t *d_order = nullptr;
    uint32_t *d_tiles = nullptr;
    uint32_t n_tiles = 0;
    std::vector<uint64_t> h_offs;
};

struct hb_query {
    hb_ctx *ctx = nullptr;
    int32_t qlen = 0, qlen_pad = 0, n_ctx = 0, window = 0, threshold = 0, diag_len = 0;
    uint8_t *d_query = nullptr;
    int32_t *d_ctx = nullptr;   // [off | xdrop | cutoff] x n_ctx
    int8_t *d_mat = nullptr;    // 32x32
    uint2 *d_cells = nullptr;
    uint32_t *d_start = nullptr;
    uint32_t *d_entries = nullptr;
    uint32_t *d_pv = nullptr;
    uint32_t n_entries = 0;
    uint32_t n_cells_used = 0;
};

struct hb_scan {
    hb_ctx *ctx = nullptr;
    hb_db *db = nullptr;
    hb_query *q = nullptr;
    hb_hsp *d_out = nullptr;
    uint32_t cap = 0;
    uint32_t *d_misc = nullptr;   // [0] HSP count, [1] tile counter
    uint32_t *d_stats = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool waited = false;
    uint32_t n_out = 0;
};

extern "C" {

const char *hb_strerror(int err)
{
    switch (err) {
    case HB_OK: return "ok";
    case HB_ERR_ARG: return "invalid argument";
    case HB_ERR_NOMEM: return "out of memory";
    case HB_ERR_DEVICE: return g_last_hip[0] ? g_last_hip : "HIP runtime error";
    case HB_ERR_CAPACITY: return "output array too small";
    case HB_ERR_RESIDUE: return "residue code >= 28";
    case HB_ERR_STATE: return "handle in wrong state";
    case HB_ERR_UNSUPPORTED: return "unsupported parameters";
    case HB_ERR_BUSY: return "busy";
    default: return "unknown error";
    }
}

const int8_t *hb_blosum62(void) { return &kBlosum62[0][0]; }

void hb_params_default(hb_params *p)
{
    if (!p) return;
    p->word_size = 3;
    p->threshold = 11;
    p->window = 40;
    p->matrix = nullptr;
}

int hb_create(int device, hb_ctx **out)
{
    if (!out) return HB_ERR_ARG;
    *out = nullptr;
    int n = 0;
    HB_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return HB_ERR_ARG;
    HB_CHECK(hipSetDevice(device));
    hb_ctx *c = new (std::nothrow) hb_ctx;
    if (!c) return HB_ERR_NOMEM;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return HB_ERR_DEVICE;
    }
    c->n_cu = prop.multiProcessorCount;
    c->max_lds = prop.sharedMemPerBlock;
    int lds_optin = 0;
    if (hipDeviceGetAttribute(&lds_optin, hipDeviceAttributeSharedMemPerBlockOptin, device) ==
            hipSuccess && (size_t)lds_optin > c->max_lds)
        c->max_lds = (size_t)lds_optin;
    *out = c;
    return HB_OK;
}

int hb_destroy(hb_ctx *c)
{
    if (!c) return HB_ERR_ARG;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &b : c->pool) hipFree(b.first);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return HB_OK;
}

int hb_mem_info(hb_ctx *c, uint64_t *free_b, uint64_t *total_b)
{
    if (!c) return HB_ERR_ARG;
    size_t f = 0, t = 0;
    HB_CHECK(hipSetDevice(c->device));
    HB_CHECK(hipMemGetInfo(&f, &t));
    if (free_b) *free_b = f;
    if (total_b) *total_b = t;
    return HB_OK;
}

int hb_synchronize(hb_ctx *c)
{
    if (!c) return HB_ERR_ARG;
    HB_CHECK(hipSetDevice(c->device));
    HB_CHECK(hipStreamSynchronize(c->stream));
    return HB_OK;
}

// ------------------------------------------------------------------ database
int hb_db_free(hb_db *db)
{
    if (!db) return HB_ERR_ARG;
    hipSetDevice(db->ctx->device);
    hipFree(db->d_res);
    hipFree(db->d_offs);
    hipFree(db->d_order);
    hipFree(db->d_tiles);
    delete db;
    return HB_OK;
}

int hb_db_upload(hb_ctx *c, const uint8_t *residues, const uint64_t *offsets, uint32_t n,
                 hb_db **out)
{
    if (!c || !offsets || !out || (n > 0 && !residues)) return HB_ERR_ARG;
    *out = nullptr;
    // subject lengths are int32 on the device and diagonal state keeps
    // position + offset below 2^31: cap a subject at 2^28 residues
    for (uint32_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i] || offsets[i + 1] - offsets[i] > (1ull << 28))
            return HB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    hb_db *db = new (std::nothrow) hb_db;
    if (!db) return HB_ERR_NOMEM;
    db->ctx = c;
    db->n = n;
    const uint64_t first = n ? offsets[0] : 0;
    db->total = n ? offsets[n] - first : 0;
    db->h_offs.resize((size_t)n + 1);
    for (uint32_t i = 0; i <= n; ++i) db->h_offs[i] = n ? offsets[i] - first : 0;

    // Length-descending processing order, then tiles of <= 64 subjects and
    // ~16K residues: one queue pop per tile keeps atomics off the hot path
    // and the longest subjects start first.
    std::vector<uint32_t> order(n);
    std::iota(order.begin(), order.end(), 0u);
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        return db->h_offs[x + 1] - db->h_offs[x] > db->h_offs[y + 1] - db->h_offs[y];
    });
    std::vector<uint32_t> tiles;
    tiles.push_back(0);
    uint64_t acc = 0;
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
        acc += db->h_offs[order[i] + 1] - db->h_offs[order[i]];
        ++cnt;
        if (cnt >= 64 || acc >= 16384) {
            tiles.push_back(i + 1);
            acc = 0;
            cnt = 0;
        }
    }
    if (tiles.back() != n) tiles.push_back(n);
    db->n_tiles = (uint32_t)tiles.size() - 1;

    int rc = HB_OK;
    uint32_t *d_flag = nullptr;
    auto fail = [&](int r) {
        hipFree(d_flag);
        hb_db_free(db);
        return r;
    };
    if (hipMalloc(&db->d_res, db->total + 64) != hipSuccess ||
        hipMalloc(&db->d_offs, ((size_t)n + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&db->d_order, (size_t)std::max<uint32_t>(n, 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&db->d_tiles, tiles.size() * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_flag, sizeof(uint32_t)) != hipSuccess)
        return fail(HB_ERR_NOMEM);
    if (hipMemcpy(db->d_res, residues + first, db->total, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(db->d_res + db->total, 0, 64) != hipSuccess ||
        hipMemcpy(db->d_offs, db->h_offs.data(), ((size_t)n + 1) * sizeof(uint64_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (n && hipMemcpy(db->d_order, order.data(), (size_t)n * sizeof(uint32_t),
                        hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(db->d_tiles, tiles.data(), tiles.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(d_flag, 0, sizeof(uint32_t)) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    if (hb_launch_check_residues(db->d_res, db->total, d_flag, c->stream) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    uint32_t flag = 0;
    if (hipMemcpyAsync(&flag, d_flag, sizeof(flag), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    if (flag) return fail(HB_ERR_RESIDUE);
    hipFree(d_flag);
    *out = db;
    return rc;
}

uint32_t hb_db_subjects(const hb_db *db) { return db ? db->n : 0; }
uint64_t hb_db_residues(const hb_db *db) { return db ? db->total : 0; }

// ------------------------------------------------------------------ query
int hb_query_free(hb_query *q)
{
    if (!q) return HB_ERR_ARG;
    hipSetDevice(q->ctx->device);
    hipFree(q->d_query);
    hipFree(q->d_ctx);
    hipFree(q->d_mat);
    hipFree(q->d_cells);
    hipFree(q->d_start);
    hipFree(q->d_entries);
    hipFree(q->d_pv);
    delete q;
    return HB_OK;
}

// per wave: diagonal state + marker, 64 segment starts, job and blocked
// queues (8 B entries), staged subject
static size_t scan_wave_lds(int32_t diag_len)
{
    return 8 * (size_t)diag_len + 256 + 8 * (size_t)(HB_JOBCAP + HB_BLKCAP) + HB_SCAP;
}
// per workgroup: score matrix, presence vector, context tables, query
static int32_t ctx_pad_of(int32_t n_ctx) { return (n_ctx + 3) & ~3; }
static size_t scan_fixed_lds(int32_t qlen_pad, int32_t n_ctx)
{
    return 1024 + 4 * HB_PV_WORDS + 12 * (size_t)ctx_pad_of(n_ctx) + qlen_pad;
}

int hb_query_prepare(hb_ctx *c, const uint8_t *seq, int32_t len, int32_t n_ctx,
                     const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                     const int32_t *cutoff, const hb_params *params, hb_query **out)
{
    if (!c || !seq || len <= 0 || n_ctx <= 0 || !ctx_off || !ctx_len || !xdrop || !cutoff || !out)
        return HB_ERR_ARG;
    *out = nullptr;
    hb_params p;
    hb_params_default(&p);
    if (params) p = *params;
    if (p.word_size != 3) return HB_ERR_UNSUPPORTED;
    // queued extensions pack the query offset in 24 bits and s - last_hit
    // (< window) in 8 bits
    if (p.window <= 3) return HB_ERR_ARG;
    if (p.window > 255 || len >= (1 << 24)) return HB_ERR_UNSUPPORTED;
    for (int32_t i = 0; i < len; ++i)
        if (seq[i] >= HB_ALPHA) return HB_ERR_RESIDUE;
    std::vector<uint8_t> qflag((size_t)len);
    for (int32_t i = 0; i < len; ++i) qflag[i] = seq[i];
    int32_t prev_end = 0;
    for (int32_t k = 0; k < n_ctx; ++k) {
        if (ctx_len[k] < 0 || ctx_off[k] < prev_end || ctx_off[k] + ctx_len[k] > len ||
            xdrop[k] <= 0)
            return HB_ERR_ARG;
        prev_end = ctx_off[k] + ctx_len[k];
        for (int32_t o = ctx_off[k]; o + 3 <= ctx_off[k] + ctx_len[k]; ++o) qflag[o] |= 0x80;
    }
    hb_query *q = new (std::nothrow) hb_query;
    if (!q) return HB_ERR_NOMEM;
    q->ctx = c;
    q->qlen = len;
    q->qlen_pad = (len + 15) & ~15;
    q->n_ctx = n_ctx;
    q->window = p.window;
    q->threshold = p.threshold;
    // Diagonals are (q - s) mod diag_len.  Any diag_len >= qlen + window
    // gives the same two-hit decisions as the CPU word finder's power-of-two
    // table (two hits that alias are always more than a window apart).
    q->diag_len = (len + p.window + 63) & ~63;
    if (scan_fixed_lds(q->qlen_pad, n_ctx) + scan_wave_lds(q->diag_len) > c->max_lds) {
        delete q;
        return HB_ERR_UNSUPPORTED;
    }
    int8_t mat32[32 * 32];
    for (int i = 0; i < 32; ++i)
        for (int j = 0; j < 32; ++j)
            mat32[i * 32 + j] = (i < HB_ALPHA && j < HB_ALPHA)
                                    ? (p.matrix ? p.matrix[i * HB_ALPHA + j] : kBlosum62[i][j])
                                    : (int8_t)-128;
    std::vector<int32_t> ctxv(3 * (size_t)n_ctx);
    for (int32_t k = 0; k < n_ctx; ++k) {
        ctxv[k] = ctx_off[k];
        ctxv[n_ctx + k] = xdrop[k];
        ctxv[2 * n_ctx + k] = cutoff[k];
    }
    std::vector<uint8_t> qpad((size_t)q->qlen_pad, 0);
    memcpy(qpad.data(), seq, (size_t)len);

    std::lock_guard<std::mutex> lk(c->mu);
    auto fail = [&](int r) {
        hb_query_free(q);
        return r;
    };
    if (hipSetDevice(c->device) != hipSuccess) return fail(HB_ERR_DEVICE);
    uint8_t *d_qflag = nullptr;
    uint32_t *d_total = nullptr;
    if (hipMalloc(&q->d_query, q->qlen_pad) != hipSuccess ||
        hipMalloc(&q->d_ctx, ctxv.size() * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&q->d_mat, sizeof(mat32)) != hipSuccess ||
        hipMalloc(&q->d_cells, HB_NCELLS * sizeof(uint2)) != hipSuccess ||
        hipMalloc(&q->d_start, (HB_NCELLS + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&q->d_pv, HB_PV_WORDS * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_qflag, (size_t)len) != hipSuccess ||
        hipMalloc(&d_total, sizeof(uint32_t)) != hipSuccess) {
        hipFree(d_qflag);
        hipFree(d_total);
        return fail(HB_ERR_NOMEM);
    }
    hipStream_t st = c->stream;
    bool ok = hipMemcpyAsync(q->d_query, qpad.data(), q->qlen_pad, hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hipMemcpyAsync(q->d_ctx, ctxv.data(), ctxv.size() * sizeof(int32_t),
                             hipMemcpyHostToDevice, st) == hipSuccess &&
              hipMemcpyAsync(q->d_mat, mat32, sizeof(mat32), hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hipMemcpyAsync(d_qflag, qflag.data(), (size_t)len, hipMemcpyHostToDevice, st) ==
                  hipSuccess &&
              hb_launch_lut_count(d_qflag, len, nullptr, q->d_mat, p.threshold, q->d_start, st) ==
                  hipSuccess &&
              hb_launch_lut_scan(q->d_start, d_total, st) == hipSuccess;
    uint32_t total = 0;
    ok = ok && hipMemcpyAsync(&total, d_total, sizeof(total), hipMemcpyDeviceToHost, st) ==
                   hipSuccess &&
         hipStreamSynchronize(st) == hipSuccess;
    if (ok) {
        q->n_entries = total;
        ok = hipMalloc(&q->d_entries, (size_t)std::max<uint32_t>(total, 1) * sizeof(uint32_t)) ==
                 hipSuccess &&
             hb_launch_lut_fill(d_qflag, len, nullptr, q->d_mat, p.threshold, q->d_start,
                                q->d_entries, q->d_cells, q->d_pv, st) == hipSuccess &&
             hipStreamSynchronize(st) == hipSuccess;
    }
    hipFree(d_qflag);
    hipFree(d_total);
    if (!ok) return fail(HB_ERR_DEVICE);
    {
        std::vector<uint32_t> pv(HB_PV_WORDS);
        if (hipMemcpy(pv.data(), q->d_pv, HB_PV_WORDS * 4, hipMemcpyDeviceToHost) != hipSuccess)
            return fail(HB_ERR_DEVICE);
        uint32_t used = 0;
        for (uint32_t w : pv) used += (uint32_t)__builtin_popcount(w);
        q->n_cells_used = used;
    }
    *out = q;
    return HB_OK;
}

int hb_query_lut_info(const hb_query *q, uint32_t *n_entries, uint32_t *n_cells_used)
{
    if (!q) return HB_ERR_ARG;
    if (n_entries) *n_entries = q->n_entries;
    if (n_cells_used) *n_cells_used = q->n_cells_used;
    return HB_OK;
}

int hb_query_lut_fetch(const hb_query *q, uint32_t *cell_start, uint32_t *entries, uint32_t cap)
{
    if (!q || !cell_start || (!entries && q->n_entries)) return HB_ERR_ARG;
    if (cap < q->n_entries) return HB_ERR_CAPACITY;
    HB_CHECK(hipSetDevice(q->ctx->device));
    HB_CHECK(hipMemcpy(cell_start, q->d_start, (HB_NCELLS + 1) * sizeof(uint32_t),
                       hipMemcpyDeviceToHost));
    if (q->n_entries)
        HB_CHECK(hipMemcpy(entries, q->d_entries, (size_t)q->n_entries * sizeof(uint32_t),
                           hipMemcpyDeviceToHost));
    return HB_OK;
}

// ------------------------------------------------------------------ scan
static int launch_scan(hb_scan *s)
{
    hb_ctx *c = s->ctx;
    hb_db *db = s->db;
    hb_query *q = s->q;
    const size_t wave_lds = scan_wave_lds(q->diag_len);
    const size_t fixed = scan_fixed_lds(q->qlen_pad, q->n_ctx);
    // 4 wavefronts per workgroup share the query tables; fewer when the
    // diagonal arrays of long (e.g. six-frame) queries do not fit.
    int waves = 4;
    while (waves > 1 && fixed + waves * wave_lds > c->max_lds) waves /= 2;
    const size_t lds = fixed + waves * wave_lds;
    if (lds > c->max_lds) return HB_ERR_UNSUPPORTED;
    int per_cu = 0;
    HB_CHECK(hb_scan_occupancy(lds, waves * HB_WAVE, &per_cu));
    if (per_cu < 1) per_cu = 1;
    int blocks = c->n_cu * per_cu;
    const int need = (int)((db->n_tiles + waves - 1) / waves);
    if (need < blocks) blocks = std::max(need, 1);

    hb_scan_args a;
    memset(&a, 0, sizeof(a));
    a.res = db->d_res;
    a.offs = db->d_offs;
    a.order = db->d_order;
    a.tiles = db->d_tiles;
    a.n_tiles = db->n_tiles;
    a.cells = q->d_cells;
    a.entries = q->d_entries;
    a.n_entries = std::max<uint32_t>(q->n_entries, 1);
    a.pv = q->d_pv;
    a.query = q->d_query;
    a.qlen = q->qlen;
    a.qlen_pad = q->qlen_pad;
    a.ctx_off = q->d_ctx;
    a.ctx_xdrop = q->d_ctx + q->n_ctx;
    a.ctx_cutoff = q->d_ctx + 2 * q->n_ctx;
    a.n_ctx = q->n_ctx;
    a.ctx_pad = ctx_pad_of(q->n_ctx);
    a.mat32 = q->d_mat;
    a.window = q->window;
    a.diag_len = q->diag_len;
    a.wave_lds = (uint32_t)wave_lds;
    a.tile_counter = s->d_misc + 1;
    a.out = s->d_out;
    a.out_cap = s->cap;
    a.out_count = s->d_misc;
    a.stats = s->d_stats;
    HB_CHECK(hipMemsetAsync(s->d_misc, 0, 2 * sizeof(uint32_t), c->stream));
    HB_CHECK(hipEventRecord(s->ev0, c->stream));
    if (db->n_tiles) HB_CHECK(hb_launch_scan(a, blocks, waves * HB_WAVE, lds, c->stream));
    HB_CHECK(hipEventRecord(s->ev1, c->stream));
    return HB_OK;
}

int hb_scan_free(hb_scan *s)
{
    if (!s) return HB_ERR_ARG;
    hipSetDevice(s->ctx->device);
    if (!s->waited) hipStreamSynchronize(s->ctx->stream);
    pool_put(s->ctx, s->d_out, (size_t)s->cap * sizeof(hb_hsp));
    pool_put(s->ctx, s->d_misc, 2 * sizeof(uint32_t));
    pool_put(s->ctx, s->d_stats, (size_t)std::max<uint32_t>(s->db->n, 1) * 3 * sizeof(uint32_t));
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    delete s;
    return HB_OK;
}

int hb_scan_start(hb_ctx *c, hb_db *db, hb_query *q, hb_scan **out)
{
    if (!c || !db || !q || !out || db->ctx != c || q->ctx != c) return HB_ERR_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    hb_scan *s = new (std::nothrow) hb_scan;
    if (!s) return HB_ERR_NOMEM;
    s->ctx = c;
    s->db = db;
    s->q = q;
    s->cap = std::max<uint32_t>(1u << 18, db->n / 4);
    s->d_out = (hb_hsp *)pool_get(c, (size_t)s->cap * sizeof(hb_hsp));
    s->d_misc = (uint32_t *)pool_get(c, 2 * sizeof(uint32_t));
    s->d_stats = (uint32_t *)pool_get(c, (size_t)std::max<uint32_t>(db->n, 1) * 3 * sizeof(uint32_t));
    if (!s->d_out || !s->d_misc || !s->d_stats) {
        hb_scan_free(s);
        return HB_ERR_NOMEM;
    }
    if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        hb_scan_free(s);
        return HB_ERR_DEVICE;
    }
    int rc = launch_scan(s);
    if (rc) {
        hb_scan_free(s);
        return rc;
    }
    *out = s;
    return HB_OK;
}

int hb_scan_wait(hb_scan *s)
{
    if (!s) return HB_ERR_ARG;
    if (s->waited) return HB_OK;
    hb_ctx *c = s->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    for (;;) {
        uint32_t n = 0;
        HB_CHECK(hipMemcpyAsync(&n, s->d_misc, sizeof(n), hipMemcpyDeviceToHost, c->stream));
        HB_CHECK(hipStreamSynchronize(c->stream));
        if (n <= s->cap) {
            s->n_out = n;
            break;
        }
        // The kernel counted every HSP but stored only cap: rerun with room
        // for all of them (the stage is deterministic).
        pool_put(c, s->d_out, (size_t)s->cap * sizeof(hb_hsp));
        s->cap = n + n / 8 + 1024;
        s->d_out = (hb_hsp *)pool_get(c, (size_t)s->cap * sizeof(hb_hsp));
        if (!s->d_out) {
            s->cap = 0;
            return HB_ERR_NOMEM;
        }
        int rc = launch_scan(s);
        if (rc) return rc;
    }
    s->waited = true;
    return HB_OK;
}

int hb_scan_elapsed_ms(hb_scan *s, float *ms)
{
    if (!s || !ms) return HB_ERR_ARG;
    int rc = hb_scan_wait(s);
    if (rc) return rc;
    HB_CHECK(hipEventElapsedTime(ms, s->ev0, s->ev1));
    return HB_OK;
}

int hb_scan_results(hb_scan *s, hb_subject_stats *stats, hb_hsp *hsps, uint64_t cap,
                    uint64_t *n_hsp)
{
    if (!s) return HB_ERR_ARG;
    int rc = hb_scan_wait(s);
    if (rc) return rc;
    const uint64_t n = s->n_out;
    if (n_hsp) *n_hsp = n;
    if (!hsps && !stats) return HB_OK;
    if (hsps && cap < n) return HB_ERR_CAPACITY;
    std::vector<hb_hsp> v((size_t)n);
    HB_CHECK(hipSetDevice(s->ctx->device));
    if (n) HB_CHECK(hipMemcpy(v.data(), s->d_out, (size_t)n * sizeof(hb_hsp), hipMemcpyDeviceToHost));
    std::sort(v.begin(), v.end(), [](const hb_hsp &x, const hb_hsp &y) {
        if (x.oid != y.oid) return x.oid < y.oid;
        if (x.s_off != y.s_off) return x.s_off < y.s_off;
        return x.q_off < y.q_off;
    });
    if (hsps && n) memcpy(hsps, v.data(), (size_t)n * sizeof(hb_hsp));
    if (stats) {
        const uint32_t ns = s->db->n;
        std::vector<uint32_t> raw((size_t)ns * 3);
        if (ns)
            HB_CHECK(hipMemcpy(raw.data(), s->d_stats, raw.size() * sizeof(uint32_t),
                               hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < ns; ++i) {
            stats[i].lookup_hits = raw[3 * (size_t)i];
            stats[i].extensions = raw[3 * (size_t)i + 1];
            stats[i].n_hsp = raw[3 * (size_t)i + 2];
            stats[i].first = -1;
        }
        for (uint64_t k = 0; k < n; ++k)
            if (k == 0 || v[k].oid != v[k - 1].oid) stats[v[k].oid].first = (int64_t)k;
    }
    return HB_OK;
}

int hb_search(hb_ctx *c, hb_db *db, const uint8_t *seq, int32_t len, int32_t n_ctx,
              const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
              const int32_t *cutoff, const hb_params *params, hb_subject_stats *stats,
              hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp)
{
    hb_query *q = nullptr;
    int rc = hb_query_prepare(c, seq, len, n_ctx, ctx_off, ctx_len, xdrop, cutoff, params, &q);
    if (rc) return rc;
    hb_scan *s = nullptr;
    rc = hb_scan_start(c, db, q, &s);
    if (!rc) rc = hb_scan_results(s, stats, hsps, cap, n_hsp);
    if (s) hb_scan_free(s);
    hb_query_free(q);
    return rc;
}

// ------------------------------------------------------------------ diagnostics
int hb_diag_copy(hb_ctx *c, uint64_t bytes, int reps, float *ms_per_copy)
{
    if (!c || bytes < 16 || reps < 1 || !ms_per_copy) return HB_ERR_ARG;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    bytes &= ~(uint64_t)15;
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = HB_OK;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        rc = HB_ERR_NOMEM;
    } else if (hipMemsetAsync(a, 1, bytes, c->stream) != hipSuccess ||
               hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
               hb_launch_copy(b, a, bytes, c->stream) != hipSuccess ||
               hipEventRecord(e0, c->stream) != hipSuccess) {
        rc = HB_ERR_DEVICE;
    } else {
        for (int r = 0; r < reps && rc == HB_OK; ++r)
            if (hb_launch_copy(b, a, bytes, c->stream) != hipSuccess) rc = HB_ERR_DEVICE;
        float ms = 0;
        if (rc == HB_OK &&
            (hipEventRecord(e1, c->stream) != hipSuccess ||
             hipEventSynchronize(e1) != hipSuccess ||
             hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = HB_ERR_DEVICE;
        *ms_per_copy = ms / (float)reps;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(a);
    hipFree(b);
    return rc;
}

// ------------------------------------------------------------------ sequences
static const char kStdaa[] = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";

int hb_encode_protein(const char *text, int32_t len, uint8_t *out)
{
    if (!text || !out || len < 0) return HB_ERR_ARG;
    for (int32_t i = 0; i < len; ++i) {
        char ch = text[i];
        if (ch >= 'a' && ch <= 'z') ch = (char)(ch - 'a' + 'A');
        const char *p = strchr(kStdaa, ch);
        if (!p || ch == 0) return HB_ERR_RESIDUE;
        out[i] = (uint8_t)(p - kStdaa);
    }
    return HB_OK;
}

// Standard genetic code (NCBI translation table 1), codons indexed
// 16*b1 + 4*b2 + b3 with T=0, C=1, A=2, G=3.
static const char kCode1[] =
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";

static int nt_index(char ch)
{
    switch (ch) {
    case 'T': case 't': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'A': case 'a': return 2;
    case 'G': case 'g': return 3;
    default: return -1;
    }
}

int32_t hb_blastx_concat_len(int32_t nt_len)
{
    if (nt_len < 0) return 0;
    int32_t total = 0;
    for (int f = 0; f < 3; ++f) total += 2 * (nt_len > f ? (nt_len - f) / 3 : 0);
    return total + 5;
}

int hb_blastx_translate(const char *nt, int32_t nt_len, uint8_t *out, int32_t *ctx_off,
                        int32_t *ctx_len)
{
    if (!nt || nt_len < 0 || !out || !ctx_off || !ctx_len) return HB_ERR_ARG;
    // fwd: plus strand; rev: reverse complement (T<->A, C<->G)
    static const int comp[4] = {2, 3, 0, 1};
    std::vector<int> fwd((size_t)nt_len), rev((size_t)nt_len);
    for (int32_t i = 0; i < nt_len; ++i) {
        fwd[i] = nt_index(nt[i]);
        const int x = nt_index(nt[nt_len - 1 - i]);
        rev[i] = x < 0 ? -1 : comp[x];
    }
    int32_t pos = 0;
    for (int fr = 0; fr < 6; ++fr) {
        const std::vector<int> &src = fr < 3 ? fwd : rev;
        const int f = fr % 3;
        const int32_t n = nt_len > f ? (nt_len - f) / 3 : 0;
        if (fr) out[pos++] = 0;  // sentinel between contexts
        ctx_off[fr] = pos;
        ctx_len[fr] = n;
        for (int32_t i = 0; i < n; ++i) {
            const int b1 = src[f + 3 * i], b2 = src[f + 3 * i + 1], b3 = src[f + 3 * i + 2];
            uint8_t aa = 21;  // X
            if (b1 >= 0 && b2 >= 0 && b3 >= 0) {
                const char ch = kCode1[16 * b1 + 4 * b2 + b3];
                aa = (uint8_t)(strchr(kStdaa, ch) - kStdaa);
            }
            out[pos++] = aa;
        }
    }
    return HB_OK;
}

}  // extern "C"
