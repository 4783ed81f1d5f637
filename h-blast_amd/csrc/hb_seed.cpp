// hb_seed.cpp -- host runtime behind include/hb_seed.h (plain per-query path)
// and the pieces the batched path shares (hb_runtime.h).
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
#include <thread>
#include <vector>

#include "hb_runtime.h"

static thread_local char g_last_hip[256];

void hb_note_hip_error(hipError_t e, const char *what, int line)
{
    snprintf(g_last_hip, sizeof(g_last_hip), "%s (%s) at line %d", hipGetErrorString(e), what,
             line);
}

// BLOSUM62 in ncbistdaa order -,A,B,C,D,E,F,G,H,I,K,L,M,N,P,Q,R,S,T,V,W,X,Y,Z,U,*,O,J.
// Same values as the reference's table (h_blast_v1.1/H-BLAST/h_blast.c:42-71);
// the gap letter scores -128 so that context sentinels stop every extension.
const int8_t kHbBlosum62[HB_ALPHA][HB_ALPHA] = {
#include "hb_blosum62.inc"
};

void *hb_pool_get(hb_ctx *c, size_t bytes)
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
            c->pool_bytes -= c->pool[best].second;
            c->pool.erase(c->pool.begin() + (long)best);
            return p;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    return p;
}

// Returns a buffer to the pool.  hipFree waits for the whole device, so the
// pool frees (oldest first) only past 64 buffers or 16 GiB.
void hb_pool_put(hb_ctx *c, void *p, size_t bytes)
{
    if (!p) return;
    std::lock_guard<std::mutex> lk(c->pool_mu);
    c->pool.emplace_back(p, bytes);
    c->pool_bytes += bytes;
    while (c->pool.size() > 64 || c->pool_bytes > ((size_t)16 << 30)) {
        hipFree(c->pool.front().first);
        c->pool_bytes -= c->pool.front().second;
        c->pool.erase(c->pool.begin());
    }
}

int hb_copy_out(hb_ctx *c, void *dst, const void *src, size_t bytes, hipStream_t st)
{
    if (bytes == 0) return HB_OK;
    std::lock_guard<std::mutex> lk(c->pin_mu);
    if (c->pinned_bytes < bytes) {
        if (c->pinned) hipHostFree(c->pinned);
        c->pinned = nullptr;
        c->pinned_bytes = 0;
        const size_t want = std::max(bytes, (size_t)64 << 20);
        HB_CHECK(hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
        c->pinned_bytes = want;
    }
    HB_CHECK(hipMemcpyAsync(c->pinned, src, bytes, hipMemcpyDeviceToHost, st));
    HB_CHECK(hipStreamSynchronize(st));
    const unsigned nt = bytes < ((size_t)4 << 20)
                            ? 1u
                            : std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    const size_t per = (bytes + nt - 1) / nt;
    auto part = [&](unsigned t) {
        const size_t b = (size_t)t * per, e = std::min(bytes, b + per);
        if (b < e) memcpy(static_cast<char *>(dst) + b, static_cast<const char *>(c->pinned) + b, e - b);
    };
    if (nt == 1) {
        part(0);
    } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(part, t);
        for (auto &x : th) x.join();
    }
    return HB_OK;
}

int hb_resolve_params(const hb_params *pp, hb_resolved_params *out)
{
    hb_params p;
    hb_params_default(&p);
    if (pp) p = *pp;
    if (p.word_size != 3) return HB_ERR_UNSUPPORTED;
    // queued extensions pack the query offset in 24 bits and s - last_hit
    // (< window) in 8 bits
    if (p.window <= 3) return HB_ERR_ARG;
    if (p.window > 255) return HB_ERR_UNSUPPORTED;
    out->threshold = p.threshold;
    out->window = p.window;
    for (int i = 0; i < 32; ++i)
        for (int j = 0; j < 32; ++j)
            out->mat32[i * 32 + j] =
                (i < HB_ALPHA && j < HB_ALPHA)
                    ? (p.matrix ? p.matrix[i * HB_ALPHA + j] : kHbBlosum62[i][j])
                    : (int8_t)-128;
    return HB_OK;
}

int hb_check_query(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                   const int32_t *ctx_len, const int32_t *xdrop)
{
    if (!seq || len <= 0 || n_ctx <= 0 || !ctx_off || !ctx_len || !xdrop) return HB_ERR_ARG;
    if (len >= (1 << 24)) return HB_ERR_UNSUPPORTED;
    for (int32_t i = 0; i < len; ++i)
        if (seq[i] >= HB_ALPHA) return HB_ERR_RESIDUE;
    int32_t prev_end = 0;
    for (int32_t k = 0; k < n_ctx; ++k) {
        if (ctx_len[k] < 0 || ctx_off[k] < prev_end || ctx_off[k] + ctx_len[k] > len ||
            xdrop[k] <= 0)
            return HB_ERR_ARG;
        prev_end = ctx_off[k] + ctx_len[k];
    }
    return HB_OK;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

void hb_group_set_free(hb_group_set *g)
{
    if (!g || !g->ctx) return;
    hb_pool_put(g->ctx, g->d_desc, g->desc.size() * sizeof(hb_group_desc));
    hb_pool_put(g->ctx, g->d_fixed, g->fixed_bytes);
    hb_pool_put(g->ctx, g->d_entries, g->entries_bytes);
    hb_pool_put(g->ctx, g->d_mat, 1024);
    g->d_desc = nullptr;
    g->d_fixed = nullptr;
    g->d_entries = nullptr;
    g->d_mat = nullptr;
}

// Lookup tables of all groups are built together: count pass, per-group
// exclusive scan, one read-back of the totals, entries allocation, fill pass.
int hb_group_set_build(hb_ctx *c, const std::vector<hb_concat> &qs,
                       const std::vector<int32_t> &ctx_base, const hb_resolved_params &p,
                       hipStream_t st, hb_group_set *out)
{
    const size_t n = qs.size();
    if (n == 0 || ctx_base.size() != n) return HB_ERR_ARG;
    out->ctx = c;
    out->desc.assign(n, hb_group_desc{});
    // per-group layout inside d_fixed
    const size_t cells_b = HB_NCELLS * sizeof(uint2);
    const size_t start_b = align256((HB_NCELLS + 1) * sizeof(uint32_t));
    const size_t pv_b = HB_PV_WORDS * sizeof(uint32_t);
    std::vector<size_t> off(n + 1, 0);
    int32_t qlen_max = 0;
    for (size_t g = 0; g < n; ++g) {
        const int32_t qlen = (int32_t)qs[g].seq.size();
        const int32_t qpad = (qlen + 15) & ~15;
        const int32_t nctx = (int32_t)qs[g].ctx_off.size();
        const size_t bytes = cells_b + start_b + pv_b + align256(qpad) +
                             align256(3 * (size_t)nctx * sizeof(int32_t)) + align256(qlen) + 256;
        off[g + 1] = off[g] + bytes;
        qlen_max = std::max(qlen_max, qlen);
        out->qlen_pad_max = std::max(out->qlen_pad_max, qpad);
        out->ctx_pad_max = std::max(out->ctx_pad_max, hb_ctx_pad(nctx));
        out->diag_len_max = std::max(out->diag_len_max, (qlen + p.window + 63) & ~63);
    }
    out->fixed_bytes = off[n];
    // host image of the inputs (query, ctx, qflag) and the job table
    std::vector<uint8_t> img(out->fixed_bytes, 0);
    std::vector<hb_lut_job> jobs(n);
    out->d_fixed = hb_pool_get(c, out->fixed_bytes);
    out->d_mat = static_cast<int8_t *>(hb_pool_get(c, 1024));
    out->d_desc = static_cast<hb_group_desc *>(hb_pool_get(c, n * sizeof(hb_group_desc)));
    if (!out->d_fixed || !out->d_mat || !out->d_desc) {
        hb_group_set_free(out);
        return HB_ERR_NOMEM;
    }
    uint8_t *base = static_cast<uint8_t *>(out->d_fixed);
    for (size_t g = 0; g < n; ++g) {
        const hb_concat &Q = qs[g];
        const int32_t qlen = (int32_t)Q.seq.size();
        const int32_t qpad = (qlen + 15) & ~15;
        const int32_t nctx = (int32_t)Q.ctx_off.size();
        size_t o = off[g];
        uint2 *cells = reinterpret_cast<uint2 *>(base + o);
        o += cells_b;
        uint32_t *start = reinterpret_cast<uint32_t *>(base + o);
        o += start_b;
        uint32_t *pv = reinterpret_cast<uint32_t *>(base + o);
        o += pv_b;
        const size_t q_o = o;
        memcpy(&img[q_o], Q.seq.data(), (size_t)qlen);
        o += align256(qpad);
        const size_t c_o = o;
        int32_t *cimg = reinterpret_cast<int32_t *>(&img[c_o]);
        for (int32_t k = 0; k < nctx; ++k) {
            cimg[k] = Q.ctx_off[k];
            cimg[nctx + k] = Q.xdrop[k];
            cimg[2 * nctx + k] = Q.cutoff[k];
        }
        o += align256(3 * (size_t)nctx * sizeof(int32_t));
        const size_t f_o = o;
        uint8_t *qf = &img[f_o];
        for (int32_t i = 0; i < qlen; ++i) qf[i] = Q.seq[i];
        for (int32_t k = 0; k < nctx; ++k)
            for (int32_t w = Q.ctx_off[k]; w + 3 <= Q.ctx_off[k] + Q.ctx_len[k]; ++w) qf[w] |= 0x80;
        // masked residues seed no word (NCBI builds its table over the
        // unmasked segments only)
        if (!Q.mask.empty())
            for (int32_t w = 0; w + 3 <= qlen; ++w)
                if (Q.mask[w] | Q.mask[w + 1] | Q.mask[w + 2]) qf[w] &= 0x7f;
        o += align256(qlen);
        uint32_t *total = reinterpret_cast<uint32_t *>(base + o);
        hb_group_desc &D = out->desc[g];
        D.cells = cells;
        D.pv = pv;
        D.query = base + q_o;
        D.ctx = reinterpret_cast<const int32_t *>(base + c_o);
        D.qlen = qlen;
        D.qlen_pad = qpad;
        D.n_ctx = nctx;
        D.diag_len = (qlen + p.window + 63) & ~63;
        D.ctx_base = ctx_base[g];
        jobs[g].qflag = base + f_o;
        jobs[g].qlen = qlen;
        jobs[g].count_start = start;
        jobs[g].cells = cells;
        jobs[g].pv = pv;
        jobs[g].total = total;
    }
    hb_lut_job *d_jobs = static_cast<hb_lut_job *>(hb_pool_get(c, n * sizeof(hb_lut_job)));
    auto fail = [&](int rc) {
        hipStreamSynchronize(st);
        hb_pool_put(c, d_jobs, n * sizeof(hb_lut_job));
        hb_group_set_free(out);
        return rc;
    };
    if (!d_jobs) return fail(HB_ERR_NOMEM);
    // inputs: copy each group's query/ctx/qflag span (the tables are written on the device)
    for (size_t g = 0; g < n; ++g) {
        // query, context table and qflag are contiguous in the group's span
        const size_t q_o = (size_t)(out->desc[g].query - base);
        const size_t end = (size_t)(jobs[g].qflag - base) + align256(jobs[g].qlen);
        if (hipMemcpyAsync(base + q_o, &img[q_o], end - q_o, hipMemcpyHostToDevice, st) !=
            hipSuccess)
            return fail(HB_ERR_DEVICE);
    }
    if (hipMemcpyAsync(out->d_mat, p.mat32, 1024, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_jobs, jobs.data(), n * sizeof(hb_lut_job), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hb_launch_lut(d_jobs, (int)n, qlen_max, out->d_mat, p.threshold, 0, st) != hipSuccess ||
        hb_launch_lut(d_jobs, (int)n, qlen_max, out->d_mat, p.threshold, 1, st) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    std::vector<uint32_t> totals(n);
    for (size_t g = 0; g < n; ++g)
        if (hipMemcpyAsync(&totals[g], jobs[g].total, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           st) != hipSuccess)
            return fail(HB_ERR_DEVICE);
    if (hipStreamSynchronize(st) != hipSuccess) return fail(HB_ERR_DEVICE);
    size_t ent_total = 0;
    for (uint32_t t : totals) ent_total += t;
    out->entries_bytes = std::max<size_t>(ent_total, 1) * sizeof(uint32_t);
    out->d_entries = static_cast<uint32_t *>(hb_pool_get(c, out->entries_bytes));
    if (!out->d_entries) return fail(HB_ERR_NOMEM);
    size_t e = 0;
    out->n_entries.resize(n);
    for (size_t g = 0; g < n; ++g) {
        jobs[g].entries = out->d_entries + e;
        out->desc[g].entries = out->d_entries + e;
        out->desc[g].n_entries = totals[g];
        out->n_entries[g] = totals[g];
        e += totals[g];
    }
    if (hipMemcpyAsync(d_jobs, jobs.data(), n * sizeof(hb_lut_job), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hb_launch_lut(d_jobs, (int)n, qlen_max, out->d_mat, p.threshold, 2, st) != hipSuccess ||
        hipMemcpyAsync(out->d_desc, out->desc.data(), n * sizeof(hb_group_desc),
                       hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    hb_pool_put(c, d_jobs, n * sizeof(hb_lut_job));
    return HB_OK;
}

// Picks wavefronts per workgroup and the staging ring size: the most
// wavefronts per CU up to 16 first (measured on C1: 8 -> 12 -> 16 wavefronts
// per CU gave 27.7 -> 20.6 -> 16.8 ms; beyond 16 the kernel is bound by
// VALU issue and more wavefronts gained nothing), then the largest ring
// (subjects above it read their residues from global memory, and a larger
// ring holds the extensions of more subjects before they must run), then
// more wavefronts per workgroup (they share the query tables).  The batched
// form stages at most 2048 residues per wave: its largest class of C3
// queries lost a workgroup per CU to larger staging.
int hb_scan_config(hb_ctx *c, bool multi, int32_t qlen_pad_max, int32_t ctx_pad_max,
                   int32_t diag_len_max, uint32_t work_items, hb_launch_cfg *cfg, int place,
                   int32_t ring_cap, int32_t force_waves)
{
    const size_t fixed = hb_scan_fixed_lds(qlen_pad_max, ctx_pad_max);
    const bool gstate = place == HB_PLACE_GLOBAL;
    const int sb = hb_state_bytes(place);
    int32_t ring_top = multi ? std::min(HB_RING_MAX, 2048) : HB_RING_MAX;
    if (ring_cap > 0) ring_top = std::min(ring_top, ring_cap & ~63);
    int best_wpc = 0;
    for (int waves : {4, 2, 1}) {
        if (force_waves > 0 && waves != force_waves) continue;
        for (int32_t ring = ring_top; ring >= 256; ring -= 64) {
            const size_t lds = fixed + waves * hb_scan_wave_lds(diag_len_max, ring, sb);
            if (lds > c->max_lds) continue;
            int per_cu = 0;
            HB_CHECK(hb_scan_occupancy(multi, gstate, lds, waves * HB_WAVE, &per_cu));
            const int wpc = std::min(per_cu * waves, 16);
            if (wpc > best_wpc || (wpc == best_wpc && ring > cfg->ring)) {
                best_wpc = wpc;
                cfg->waves = waves;
                cfg->lds = lds;
                cfg->ring = ring;
                cfg->blocks = per_cu;
            }
        }
    }
    if (best_wpc == 0) return HB_ERR_UNSUPPORTED;
    int blocks = c->n_cu * std::max(cfg->blocks, 1);
    if (gstate) {   // keep the global-memory state within HB_GSTATE_MAX_BYTES
        const size_t per_block = HB_GSTATE_WORD * (size_t)diag_len_max * (size_t)cfg->waves;
        blocks = std::min(blocks, (int)std::max<size_t>(1, HB_GSTATE_MAX_BYTES / per_block));
    }
    const uint32_t per_block = multi ? 1u : (uint32_t)cfg->waves;
    const int need = (int)((work_items + per_block - 1) / per_block);
    if (need < blocks) blocks = std::max(need, 1);
    cfg->blocks = blocks;
    return HB_OK;
}

// Automatic placement of the diagonal state.  LDS words beat the
// global-memory words (ordering markers in an LDS table, one load and one
// store per hit) down to about kGstateMinWaves wavefronts per CU (measured
// on 200,000 subjects, ms LDS / global: 1,000 residues 1.61 / 1.84, 8,000 at
// 4 wavefronts per CU 23.7 / 27.6, 12,000 at 3 64.7 / 39.0, 31,000 at 1
// 320 / 114).
static constexpr int kGstateMinWaves = 4;
static int lds_waves(const hb_ctx *c, int32_t D, size_t fixed, int state_bytes)
{
    const size_t per_wave = hb_scan_wave_lds(D, 256, state_bytes);
    if (fixed + per_wave > c->max_lds) return 0;
    // waves of one workgroup share the fixed part; up to 4 per workgroup
    return (int)std::min<size_t>((c->max_lds - fixed) / per_wave, 16);
}

int hb_state_choose(const hb_ctx *c, int32_t D, size_t fixed, uint64_t max_len, int32_t window)
{
    if (hb_state_lds_ok(max_len, window) && lds_waves(c, D, fixed, 4) >= kGstateMinWaves)
        return HB_PLACE_LDS;
    return HB_PLACE_GLOBAL;
}

bool hb_query_lds_ok(const hb_ctx *c, int32_t qlen, int32_t n_ctx)
{
    const size_t qpad = (size_t)((qlen + 15) & ~15);
    const size_t fixed = hb_scan_fixed_lds((int32_t)qpad, hb_ctx_pad(n_ctx));
    return fixed + hb_scan_wave_lds(0, 256, 0) <= c->max_lds && 1024 + qpad <= c->max_lds;
}

size_t hb_gstate_bytes_max(const hb_ctx *c, int32_t D)
{
    const size_t full = (size_t)c->n_cu * 16 * HB_GSTATE_WORD * (size_t)D;   // 16 wavefronts per CU
    return std::min(full, std::max(HB_GSTATE_MAX_BYTES, 4 * HB_GSTATE_WORD * (size_t)D));
}

struct hb_query {
    hb_ctx *ctx = nullptr;
    hb_resolved_params p;
    hb_group_set gs;
    uint32_t n_cells_used = 0;
    std::vector<int32_t> ctx_off, ctx_len;   // host copies (gapped stage)
    int mode = HB_STATE_AUTO;   // requested placement of the diagonal state
    bool lds_fits = true;       // the LDS layout can hold the diagonal words
    int32_t reset_interval = 0; // hb_query_set_reset_interval (0: the placement's limit)
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
    float gap_ms = 0.f;        // device time of the last gapped stage
    uint32_t gap_wide = 0;     // subjects it ran in the wide pass
    uint32_t *d_gstate = nullptr;   // per-wavefront diagonal words (long queries)
    size_t gstate_bytes = 0;
    int place = HB_PLACE_LDS;      // where this scan keeps the diagonal state
};

// ------------------------------------------------------------------ no-alloc scan
// Scratch of hb_scan_noalloc, one copy per GPU, held until the process ends:
// the scan's HSP cursor and tile counter (the first 256 bytes), then the
// diagonal words of a query whose state lives in global memory.  Host code
// uses the device address hipGetSymbolAddress returns (hb_ctx::scratch_dev),
// never the array's name.
#define HB_NOALLOC_HEAD 256
static __device__ __align__(256) unsigned char scratch[HB_NOALLOC_HEAD + HB_NOALLOC_STATE_BYTES];
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[256];   // per device

// The device address of this GPU's scratch (the device is current).
static int scratch_address(int device, unsigned char **out)
{
    if (device < 0 || device >= 256) return HB_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        HB_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[device];
    return HB_OK;
}


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
    case HB_ERR_FORMAT: return "malformed .hDb volume or .hDbI index";
    default: return "unknown error";
    }
}

const int8_t *hb_blosum62(void) { return &kHbBlosum62[0][0]; }

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
    // one dynamic-LDS limit for every launch of the scan kernels, so that
    // launches of different sizes on different streams never disagree; the
    // no-alloc scan's scratch address of this device
    if (hb_scan_allow_lds(c->max_lds) != hipSuccess ||
        scratch_address(device, &c->scratch_dev) != HB_OK) {
        hipStreamDestroy(c->stream);
        delete c;
        return HB_ERR_DEVICE;
    }
    *out = c;
    return HB_OK;
}

int hb_destroy(hb_ctx *c)
{
    if (!c) return HB_ERR_ARG;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (auto &b : c->pool) hipFree(b.first);
    if (c->pinned) hipHostFree(c->pinned);
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

int hb_host_alloc(hb_ctx *c, uint64_t bytes, void **out)
{
    if (!c || !out) return HB_ERR_ARG;
    *out = nullptr;
    HB_CHECK(hipSetDevice(c->device));
    if (hipHostMalloc(out, std::max<uint64_t>(bytes, 1), hipHostMallocDefault) != hipSuccess) {
        *out = nullptr;
        return HB_ERR_NOMEM;
    }
    return HB_OK;
}

int hb_host_free(hb_ctx *c, void *p)
{
    if (!c) return HB_ERR_ARG;
    if (!p) return HB_OK;
    HB_CHECK(hipSetDevice(c->device));
    HB_CHECK(hipHostFree(p));
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
    hipFree(db->d_subj);
    hipFree(db->d_scan_res);
    hipFree(db->d_tiles);
    delete db;
    return HB_OK;
}

// Processing order and tiles of a database whose host offsets are set, and
// the device copies of offsets, per-subject records in processing order and
// tiles.
static int db_meta(hb_db *db)
{
    const uint32_t n = db->n;
    // Length-descending processing order, then tiles of <= 64 subjects and
    // ~16K residues: one queue pop per tile keeps atomics off the hot path
    // and the longest subjects start first.
    // (stable LSD radix sort on the length, 11 bits a pass: a comparison
    // sort reading scattered offsets took ~0.2 s for 2M subjects)
    std::vector<uint32_t> order(n), tmp(n), key(n), key2(n);
    uint32_t kmax = 0;
    for (uint32_t i = 0; i < n; ++i) {
        order[i] = i;
        key[i] = (uint32_t)(db->h_offs[i + 1] - db->h_offs[i]);
        kmax = std::max(kmax, key[i]);
    }
    db->max_len = kmax;
    for (uint32_t i = 0; i < n; ++i) key[i] = kmax - key[i];   // ascending key = descending length
    for (int shift = 0; shift < 32 && (kmax >> shift) != 0; shift += 11) {
        uint32_t cnt[2049] = {0};
        for (uint32_t i = 0; i < n; ++i) ++cnt[((key[i] >> shift) & 2047u) + 1];
        for (int b = 0; b < 2048; ++b) cnt[b + 1] += cnt[b];
        for (uint32_t i = 0; i < n; ++i) {
            const uint32_t d = cnt[(key[i] >> shift) & 2047u]++;
            tmp[d] = order[i];
            key2[d] = key[i];
        }
        order.swap(tmp);
        key.swap(key2);
    }
    // the scan reads each subject's oid, length and offset from one 16-byte
    // record in processing order (sequential), not offs[oid] (one scattered
    // cache line per subject); the offset points into the scan's copy of
    // the residues, laid out in the same order, so consecutive subjects share
    // no cache line with subjects read at other times
    std::vector<uint4> subj(n);
    std::vector<uint32_t> tiles;
    tiles.push_back(0);
    uint64_t acc = 0, dst = 0;
    uint32_t cnt = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const uint64_t len = db->h_offs[order[i] + 1] - db->h_offs[order[i]];
        subj[i] = make_uint4(order[i], (uint32_t)len, (uint32_t)dst, (uint32_t)(dst >> 32));
        dst += (len + 15) & ~15ull;   // the scan stages subjects with 16-byte loads
        acc += len;
        ++cnt;
        if (cnt >= 64 || acc >= 16384) {
            tiles.push_back(i + 1);
            acc = 0;
            cnt = 0;
        }
    }
    if (tiles.back() != n) tiles.push_back(n);
    db->n_tiles = (uint32_t)tiles.size() - 1;
    db->scan_bytes = dst;
    if (hipMalloc(&db->d_offs, ((size_t)n + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&db->d_subj, (size_t)std::max<uint32_t>(n, 1) * sizeof(uint4)) != hipSuccess ||
        hipMalloc(&db->d_tiles, tiles.size() * sizeof(uint32_t)) != hipSuccess)
        return HB_ERR_NOMEM;
    if (hipMemcpy(db->d_offs, db->h_offs.data(), ((size_t)n + 1) * sizeof(uint64_t),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (n && hipMemcpy(db->d_subj, subj.data(), (size_t)n * sizeof(uint4),
                        hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(db->d_tiles, tiles.data(), tiles.size() * sizeof(uint32_t),
                  hipMemcpyHostToDevice) != hipSuccess)
        return HB_ERR_DEVICE;
    return HB_OK;
}

// The scan's copy of the residues in processing order (after d_res holds
// the database).
static int db_scan_copy(hb_db *db)
{
    hb_ctx *c = db->ctx;
    // padding bytes between subject spans read as 0 (the staged copies hold
    // them past each subject's end)
    if (hipMalloc(&db->d_scan_res, db->scan_bytes + 64) != hipSuccess) return HB_ERR_NOMEM;
    if (hipMemsetAsync(db->d_scan_res, 0, db->scan_bytes + 64, c->stream) != hipSuccess ||
        hb_launch_gather_subjects(db->d_res, db->d_offs, db->d_subj, db->d_tiles, db->n_tiles,
                                  db->d_scan_res, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return HB_ERR_DEVICE;
    return HB_OK;
}

// Every residue code of the uploaded database is < 28.
static int db_check_residues(hb_db *db)
{
    hb_ctx *c = db->ctx;
    uint32_t *d_flag = nullptr;
    if (hipMalloc(&d_flag, sizeof(uint32_t)) != hipSuccess) return HB_ERR_NOMEM;
    uint32_t flag = 0;
    const bool ok = hipMemsetAsync(d_flag, 0, sizeof(uint32_t), c->stream) == hipSuccess &&
                    hb_launch_check_residues(db->d_res, db->total, d_flag, c->stream) ==
                        hipSuccess &&
                    hipMemcpyAsync(&flag, d_flag, sizeof(flag), hipMemcpyDeviceToHost,
                                   c->stream) == hipSuccess &&
                    hipStreamSynchronize(c->stream) == hipSuccess;
    hipFree(d_flag);
    if (!ok) return HB_ERR_DEVICE;
    return flag ? HB_ERR_RESIDUE : HB_OK;
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
    auto fail = [&](int r) {
        hb_db_free(db);
        return r;
    };
    if (hipMalloc(&db->d_res, db->total + 64) != hipSuccess) return fail(HB_ERR_NOMEM);
    if (hipMemcpy(db->d_res, residues + first, db->total, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(db->d_res + db->total, 0, 64) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    int rc = db_meta(db);
    if (rc == HB_OK) rc = db_check_residues(db);
    if (rc == HB_OK) rc = db_scan_copy(db);
    if (rc != HB_OK) return fail(rc);
    *out = db;
    return HB_OK;
}

int hb_db_upload_hdb(hb_ctx *c, const uint8_t *hdb, uint64_t bytes, int32_t blocksx,
                     int32_t threadsx, int32_t n_seqs, const int32_t *group_len,
                     int32_t n_groups, hb_db **out)
{
    if (!c || !out || n_seqs < 0 || n_groups < 0 || blocksx < 1 || blocksx > 65536 ||
        threadsx < 1 || threadsx > 65536 || (n_groups > 0 && !group_len) || (bytes > 0 && !hdb))
        return HB_ERR_ARG;
    *out = nullptr;
    // the layout the volume must have (ncbi_blast_files/functions.cpp:844-911)
    const uint64_t T = (uint64_t)threadsx, G = (uint64_t)blocksx * T;
    if ((uint64_t)n_groups != ((uint64_t)n_seqs + G - 1) / G) return HB_ERR_FORMAT;
    std::vector<uint64_t> gbase((size_t)n_groups + 1, 0);
    uint64_t acc = 0;
    for (int32_t g = 0; g < n_groups; ++g) {
        const int32_t L = group_len[g];
        if (L < 0 || L % 16 || L > (1 << 28)) return HB_ERR_FORMAT;
        const uint64_t members = std::min<uint64_t>(G, (uint64_t)n_seqs - (uint64_t)g * G);
        gbase[g] = acc;
        acc += (members + T - 1) / T * T * (uint64_t)L;
    }
    if (acc != bytes) return HB_ERR_FORMAT;

    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    hb_db *db = new (std::nothrow) hb_db;
    if (!db) return HB_ERR_NOMEM;
    db->ctx = c;
    db->n = (uint32_t)n_seqs;
    uint8_t *d_hdb = nullptr;
    uint64_t *d_gbase = nullptr;
    int32_t *d_glen = nullptr;
    uint32_t *d_lens = nullptr, *d_bad = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    auto release = [&]() {
        hipFree(d_hdb);
        hipFree(d_gbase);
        hipFree(d_glen);
        hipFree(d_lens);
        hipFree(d_bad);
        for (hipEvent_t &e : ev)
            if (e) hipEventDestroy(e);
    };
    auto fail = [&](int r) {
        hipStreamSynchronize(c->stream);
        release();
        hb_db_free(db);
        return r;
    };
    for (hipEvent_t &e : ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(HB_ERR_DEVICE);
    const size_t ng1 = (size_t)std::max(n_groups, 1);
    if (hipMalloc(&d_hdb, std::max<uint64_t>(bytes, 4)) != hipSuccess ||
        hipMalloc(&d_gbase, ng1 * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&d_glen, ng1 * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&d_lens, (size_t)std::max(n_seqs, 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&d_bad, sizeof(uint32_t)) != hipSuccess)
        return fail(HB_ERR_NOMEM);
    if ((bytes && hipMemcpy(d_hdb, hdb, bytes, hipMemcpyHostToDevice) != hipSuccess) ||
        (n_groups && (hipMemcpy(d_gbase, gbase.data(), n_groups * sizeof(uint64_t),
                                hipMemcpyHostToDevice) != hipSuccess ||
                      hipMemcpy(d_glen, group_len, n_groups * sizeof(int32_t),
                                hipMemcpyHostToDevice) != hipSuccess)))
        return fail(HB_ERR_DEVICE);
    hb_hdb_args a{};
    a.hdb = d_hdb;
    a.group_base = d_gbase;
    a.group_len = d_glen;
    a.n_seqs = (uint32_t)n_seqs;
    a.blocksx = (uint32_t)blocksx;
    a.threadsx = (uint32_t)threadsx;
    a.lens = d_lens;
    std::vector<uint32_t> lens((size_t)n_seqs);
    if (hipEventRecord(ev[0], c->stream) != hipSuccess ||
        hb_launch_hdb_lengths(a, c->stream) != hipSuccess ||
        hipEventRecord(ev[1], c->stream) != hipSuccess ||
        (n_seqs && hipMemcpyAsync(lens.data(), d_lens, lens.size() * sizeof(uint32_t),
                                  hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    db->h_offs.resize((size_t)n_seqs + 1);
    db->h_offs[0] = 0;
    for (int32_t i = 0; i < n_seqs; ++i) db->h_offs[i + 1] = db->h_offs[i] + lens[i];
    db->total = db->h_offs[n_seqs];
    if (hipMalloc(&db->d_res, db->total + 64) != hipSuccess) return fail(HB_ERR_NOMEM);
    int rc = db_meta(db);
    if (rc != HB_OK) return fail(rc);
    a.offs = db->d_offs;
    a.res = db->d_res;
    a.bad = d_bad;
    uint32_t bad = 0;
    if (hipMemsetAsync(db->d_res + db->total, 0, 64, c->stream) != hipSuccess ||
        hipMemsetAsync(d_bad, 0, sizeof(uint32_t), c->stream) != hipSuccess ||
        hipEventRecord(ev[2], c->stream) != hipSuccess ||
        hb_launch_hdb_copy(a, c->stream) != hipSuccess ||
        hipEventRecord(ev[3], c->stream) != hipSuccess ||
        hipMemcpyAsync(&bad, d_bad, sizeof(bad), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return fail(HB_ERR_DEVICE);
    hipEventElapsedTime(&db->hdb_len_ms, ev[0], ev[1]);
    hipEventElapsedTime(&db->hdb_copy_ms, ev[2], ev[3]);
    if (bad) return fail(HB_ERR_RESIDUE);   // checked by the copy kernel
    // the volume's bytes are no longer needed: free them before the scan's
    // copy is allocated, so the upload holds at most the file plus the
    // residues, then the residues twice
    release();
    d_hdb = nullptr;
    d_gbase = nullptr;
    d_glen = nullptr;
    d_lens = nullptr;
    d_bad = nullptr;
    for (hipEvent_t &e : ev) e = nullptr;
    const int rc2 = db_scan_copy(db);
    if (rc2 != HB_OK) return fail(rc2);
    *out = db;
    return HB_OK;
}

int hb_db_fetch(const hb_db *db, uint8_t *residues, uint64_t res_cap, uint64_t *offsets)
{
    if (!db || !offsets || (db->total > 0 && !residues)) return HB_ERR_ARG;
    if (res_cap < db->total) return HB_ERR_CAPACITY;
    hb_ctx *c = db->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    std::memcpy(offsets, db->h_offs.data(), db->h_offs.size() * sizeof(uint64_t));
    if (db->total && hipMemcpy(residues, db->d_res, db->total, hipMemcpyDeviceToHost) != hipSuccess)
        return HB_ERR_DEVICE;
    return HB_OK;
}

int hb_db_hdb_info(const hb_db *db, float *lengths_ms, float *copy_ms)
{
    if (!db || !lengths_ms || !copy_ms) return HB_ERR_ARG;
    *lengths_ms = db->hdb_len_ms;
    *copy_ms = db->hdb_copy_ms;
    return HB_OK;
}

uint32_t hb_db_subjects(const hb_db *db) { return db ? db->n : 0; }
uint64_t hb_db_residues(const hb_db *db) { return db ? db->total : 0; }

// ------------------------------------------------------------------ query
int hb_query_free(hb_query *q)
{
    if (!q) return HB_ERR_ARG;
    hipSetDevice(q->ctx->device);
    // scans of this query run on the context stream; the tables go back to
    // the pool, so they must not be reused while one still reads them
    hipStreamSynchronize(q->ctx->stream);
    hb_group_set_free(&q->gs);
    delete q;
    return HB_OK;
}

int hb_query_prepare(hb_ctx *c, const uint8_t *seq, int32_t len, int32_t n_ctx,
                     const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                     const int32_t *cutoff, const hb_params *params, hb_query **out)
{
    return hb_query_prepare_masked(c, seq, len, n_ctx, ctx_off, ctx_len, xdrop, cutoff, params,
                                   nullptr, out);
}

int hb_query_prepare_masked(hb_ctx *c, const uint8_t *seq, int32_t len, int32_t n_ctx,
                            const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                            const int32_t *cutoff, const hb_params *params,
                            const uint8_t *lut_mask, hb_query **out)
{
    if (!c || !out || !cutoff) return HB_ERR_ARG;
    *out = nullptr;
    int rc = hb_check_query(seq, len, n_ctx, ctx_off, ctx_len, xdrop);
    if (rc) return rc;
    hb_query *q = new (std::nothrow) hb_query;
    if (!q) return HB_ERR_NOMEM;
    q->ctx = c;
    rc = hb_resolve_params(params, &q->p);
    if (rc) {
        delete q;
        return rc;
    }
    q->ctx_off.assign(ctx_off, ctx_off + n_ctx);
    q->ctx_len.assign(ctx_len, ctx_len + n_ctx);
    std::vector<hb_concat> qs(1);
    qs[0].seq.assign(seq, seq + len);
    qs[0].ctx_off.assign(ctx_off, ctx_off + n_ctx);
    qs[0].ctx_len.assign(ctx_len, ctx_len + n_ctx);
    qs[0].xdrop.assign(xdrop, xdrop + n_ctx);
    qs[0].cutoff.assign(cutoff, cutoff + n_ctx);
    if (lut_mask) qs[0].mask.assign(lut_mask, lut_mask + len);
    // The diagonal state (4 bytes per diagonal) stays in LDS when it fits;
    // longer queries (the reference accepts 32,768 residues,
    // H-BLAST/h_blast.cu:1293) keep it in global memory.  The query itself,
    // its context tables and the matrix are always in LDS, and so is the
    // lookup-table build's copy of the query.
    const int32_t D = (len + q->p.window + 63) & ~63;
    if (!hb_query_lds_ok(c, len, n_ctx)) {
        delete q;
        return HB_ERR_UNSUPPORTED;
    }
    const size_t fixed = hb_scan_fixed_lds((len + 15) & ~15, hb_ctx_pad(n_ctx));
    q->lds_fits = fixed + hb_scan_wave_lds(D, 256, 4) <= c->max_lds;
    std::lock_guard<std::mutex> lk(c->mu);
    if (hipSetDevice(c->device) != hipSuccess) {
        delete q;
        return HB_ERR_DEVICE;
    }
    rc = hb_group_set_build(c, qs, std::vector<int32_t>(1, 0), q->p, c->stream, &q->gs);
    if (rc) {
        delete q;
        return rc;
    }
    std::vector<uint32_t> pv(HB_PV_WORDS);
    if (hipMemcpy(pv.data(), q->gs.desc[0].pv, HB_PV_WORDS * 4, hipMemcpyDeviceToHost) !=
        hipSuccess) {
        hb_query_free(q);
        return HB_ERR_DEVICE;
    }
    for (uint32_t w : pv) q->n_cells_used += (uint32_t)__builtin_popcount(w);
    *out = q;
    return HB_OK;
}

int hb_query_set_state_mode(hb_query *q, int mode)
{
    if (!q) return HB_ERR_ARG;
    switch (mode) {
    case HB_STATE_AUTO:
    case HB_STATE_GLOBAL: q->mode = mode; return HB_OK;
    case HB_STATE_LDS:
    case HB_STATE_LDS16:   // one LDS format now; the old 16-bit request means LDS
        if (!q->lds_fits) return HB_ERR_UNSUPPORTED;
        q->mode = HB_STATE_LDS;
        return HB_OK;
    default: return HB_ERR_ARG;
    }
}

int hb_query_set_reset_interval(hb_query *q, int32_t interval)
{
    if (!q || interval < 0) return HB_ERR_ARG;
    q->reset_interval = interval;
    return HB_OK;
}

int hb_query_state_mode(const hb_query *q)
{
    if (!q) return HB_ERR_ARG;
    return q->mode;
}

// The placement a scan of q over db uses.
static int state_place(const hb_query *q, const hb_db *db)
{
    switch (q->mode) {
    case HB_STATE_LDS:
        return hb_state_lds_ok(db->max_len, q->p.window) ? HB_PLACE_LDS : -1;
    case HB_STATE_GLOBAL: return HB_PLACE_GLOBAL;
    default:
        return hb_state_choose(q->ctx, q->gs.diag_len_max,
                               hb_scan_fixed_lds(q->gs.qlen_pad_max, q->gs.ctx_pad_max),
                               db->max_len, q->p.window);
    }
}

int hb_query_lut_info(const hb_query *q, uint32_t *n_entries, uint32_t *n_cells_used)
{
    if (!q) return HB_ERR_ARG;
    if (n_entries) *n_entries = q->gs.n_entries[0];
    if (n_cells_used) *n_cells_used = q->n_cells_used;
    return HB_OK;
}

int hb_query_lut_fetch(const hb_query *q, uint32_t *cell_start, uint32_t *entries, uint32_t cap)
{
    if (!q || !cell_start) return HB_ERR_ARG;
    const uint32_t ne = q->gs.n_entries[0];
    if (!entries && ne) return HB_ERR_ARG;
    if (cap < ne) return HB_ERR_CAPACITY;
    HB_CHECK(hipSetDevice(q->ctx->device));
    std::vector<uint2> cells(HB_NCELLS);
    HB_CHECK(hipMemcpy(cells.data(), q->gs.desc[0].cells, HB_NCELLS * sizeof(uint2),
                       hipMemcpyDeviceToHost));
    // every cell holds its exclusive start, empty ones included (CSR)
    for (int i = 0; i < HB_NCELLS; ++i) cell_start[i] = cells[i].x;
    cell_start[HB_NCELLS] = ne;
    if (ne)
        HB_CHECK(hipMemcpy(entries, q->gs.desc[0].entries, (size_t)ne * sizeof(uint32_t),
                           hipMemcpyDeviceToHost));
    return HB_OK;
}

// ------------------------------------------------------------------ scan
// Kernel arguments of a plain scan of q over db with launch configuration
// cfg and placement place (outputs, counters and state filled by the caller).
static int plain_scan_args(const hb_db *db, const hb_query *q, int place,
                           const hb_launch_cfg &cfg, hb_scan_args *a)
{
    memset(a, 0, sizeof(*a));
    a->res = db->d_scan_res;   // subjects in processing order (a->subj's offsets)
    a->offs = db->d_offs;
    a->subj = db->d_subj;
    a->tiles = db->d_tiles;
    a->tile_begin = 0;
    a->tile_end = db->n_tiles;
    a->groups = q->gs.d_desc;
    a->n_groups = 1;
    a->qlen_pad_max = q->gs.qlen_pad_max;
    a->ctx_pad_max = q->gs.ctx_pad_max;
    a->diag_len_max = q->gs.diag_len_max;
    a->wave_lds = (uint32_t)hb_scan_wave_lds(q->gs.diag_len_max, cfg.ring, hb_state_bytes(place));
    a->ring = (uint32_t)cfg.ring;
    a->state_limit = hb_state_limit(place, db->max_len);
    if (q->reset_interval > 0) {
        // a subject and its successor must fit below the interval
        if ((uint64_t)q->reset_interval <= 2 * (db->max_len + (uint64_t)q->p.window) + 64)
            return HB_ERR_UNSUPPORTED;
        a->state_limit = std::min(a->state_limit, q->reset_interval - (int32_t)db->max_len);
    }
    a->mat32 = q->gs.d_mat;
    a->window = q->p.window;
    return HB_OK;
}

static int launch_scan(hb_scan *s)
{
    hb_ctx *c = s->ctx;
    hb_db *db = s->db;
    hb_query *q = s->q;
    hb_launch_cfg cfg;
    const int place = state_place(q, db);
    if (place < 0) return HB_ERR_UNSUPPORTED;   // LDS state asked for, subjects too long
    s->place = place;
    int rc = hb_scan_config(c, false, q->gs.qlen_pad_max, q->gs.ctx_pad_max, q->gs.diag_len_max,
                            db->n_tiles, &cfg, place);
    if (rc) return rc;
    if (place == HB_PLACE_GLOBAL && !s->d_gstate) {
        s->gstate_bytes =
            (size_t)cfg.blocks * cfg.waves * (size_t)q->gs.diag_len_max * HB_GSTATE_WORD;
        s->d_gstate = static_cast<uint32_t *>(hb_pool_get(c, s->gstate_bytes));
        if (!s->d_gstate) return HB_ERR_NOMEM;
    }
    hb_scan_args a;
    rc = plain_scan_args(db, q, place, cfg, &a);
    if (rc) return rc;
    a.gstate = place == HB_PLACE_GLOBAL ? s->d_gstate : nullptr;   // cleared by each wavefront
    a.work_counter = s->d_misc + 1;
    a.out = s->d_out;
    a.out_cap = s->cap;
    a.out_count = s->d_misc;
    a.stats = s->d_stats;
    a.group_totals = nullptr;
    HB_CHECK(hipMemsetAsync(s->d_misc, 0, 2 * sizeof(uint32_t), c->stream));
    HB_CHECK(hipEventRecord(s->ev0, c->stream));
    if (db->n_tiles) HB_CHECK(hb_launch_scan(a, false, cfg.blocks, cfg.waves * HB_WAVE, cfg.lds, c->stream));
    HB_CHECK(hipEventRecord(s->ev1, c->stream));
    return HB_OK;
}

// ------------------------------------------------------------------ memory fit
// Device bytes of a plain scan of q over n subjects with `residues`
// residues (0: bounded by file_bytes), loaded from a .hDb of file_bytes
// (0: packed arrays), subjects up to max_len residues:
//   upload:   the file, the residues, lengths (4 B), offsets (8 B) and
//             processing-order records (16 B) per subject, tiles;
//   resident: the residues, their processing-order copy (spans padded to
//             16 B) and the metadata;
//   scan:     the HSP reservation (32 B records), counters (12 B per
//             subject), the diagonal state in global memory if placed there;
//   results:  per-subject stats (24 B), the ordered records, the layout's
//             scratch (4 B per subject) and the list of large subjects.
static uint64_t scan_need_bytes(const hb_ctx *c, const hb_query *q, uint64_t file_bytes,
                                uint64_t residues, uint32_t n, uint64_t max_len)
{
    const uint64_t R = residues ? residues : file_bytes;
    const uint64_t meta = 8ull * (n + 1) + 16ull * n + 4ull * (n / 64 + 2);
    const uint64_t upload = file_bytes + (R + 64) + 4ull * n + meta;
    const uint64_t resident = (R + 64) + (R + 16ull * n + 64) + meta;
    const uint64_t cap = std::max<uint64_t>(1u << 18, n / 4);
    uint64_t scan = 32 * cap + 12ull * n + 64;
    int place = HB_PLACE_LDS;
    switch (q->mode) {
    case HB_STATE_GLOBAL: place = HB_PLACE_GLOBAL; break;
    case HB_STATE_LDS: place = HB_PLACE_LDS; break;
    default:
        place = hb_state_choose(c, q->gs.diag_len_max,
                                hb_scan_fixed_lds(q->gs.qlen_pad_max, q->gs.ctx_pad_max), max_len,
                                q->p.window);
    }
    if (place == HB_PLACE_GLOBAL) scan += hb_gstate_bytes_max(c, q->gs.diag_len_max);
    const uint64_t results = 24ull * n + 32 * cap + 4ull * (n + n / 4096 + 1) +
                             4096 * sizeof(hb_fin_big);
    return std::max(upload, resident + scan + results);
}

int hb_check_memory(hb_ctx *c, const hb_query *q, uint64_t file_bytes, uint64_t residues,
                    uint32_t n_seqs, uint64_t max_len, int64_t budget, uint64_t *need,
                    int32_t *fits)
{
    if (!c || !q || q->ctx != c || !need || !fits) return HB_ERR_ARG;
    *need = scan_need_bytes(c, q, file_bytes, residues, n_seqs, max_len);
    uint64_t room = (uint64_t)std::max<int64_t>(budget, 0);
    if (budget < 0) {
        uint64_t total = 0;
        int rc = hb_mem_info(c, &room, &total);
        if (rc) return rc;
    }
    *fits = *need <= room ? 1 : 0;
    return HB_OK;
}

// ------------------------------------------------------------------ co-scheduled search
// The library's device steps of the co-scheduled loop (hb_cosched.cpp).
namespace {
struct CoschedDevice {
    hb_ctx *ctx = nullptr;
    hb_query *q = nullptr;
    int64_t budget = -1;
    hb_gpu_results_fn gpu = nullptr;
    void *user = nullptr;
    hb_db *db = nullptr;
    hb_scan *scan = nullptr;
    std::vector<hb_subject_stats> stats;
    std::vector<hb_hsp> hsps;

    void drop()
    {
        if (scan) hb_scan_free(scan);
        if (db) hb_db_free(db);
        scan = nullptr;
        db = nullptr;
    }
    static CoschedDevice *self(void *u) { return static_cast<CoschedDevice *>(u); }
    static int fits(void *u, int32_t, const hb_hdb_share *sh, int32_t max_len, uint64_t *need,
                    int32_t *ok)
    {
        CoschedDevice *d = self(u);
        return hb_check_memory(d->ctx, d->q, sh->bytes, 0, (uint32_t)sh->gpu_seqs,
                               (uint64_t)std::max(max_len, 0), d->budget, need, ok);
    }
    static int launch(void *u, int32_t, const uint8_t *hdb, const hb_hdb_share *sh,
                      int32_t blocksx, int32_t threadsx, const int32_t *group_len)
    {
        CoschedDevice *d = self(u);
        int rc = hb_db_upload_hdb(d->ctx, hdb, sh->bytes, blocksx, threadsx, sh->gpu_seqs,
                                  group_len, sh->n_groups, &d->db);
        if (rc) return rc;
        rc = hb_scan_start(d->ctx, d->db, d->q, &d->scan);
        if (rc) d->drop();
        return rc;
    }
    static int fetch(void *u, int32_t volume, int64_t oid_base, double *gpu_seconds,
                     uint64_t *n_hsp)
    {
        CoschedDevice *d = self(u);
        uint64_t n = 0;
        int rc = hb_scan_results(d->scan, nullptr, nullptr, 0, &n);
        const uint32_t ns = d->db ? d->db->n : 0;
        if (!rc) {
            d->stats.resize(std::max<uint32_t>(ns, 1));
            d->hsps.resize(std::max<uint64_t>(n, 1));
            rc = hb_scan_results(d->scan, d->stats.data(), d->hsps.data(), n, &n);
        }
        float ms = 0.f;
        if (!rc) rc = hb_scan_elapsed_ms(d->scan, &ms);
        d->drop();
        if (rc) return rc;
        for (uint64_t i = 0; i < n; ++i) d->hsps[i].oid += (uint32_t)oid_base;
        *gpu_seconds = ms / 1000.0;
        *n_hsp = n;
        if (d->gpu && d->gpu(d->user, volume, oid_base, d->stats.data(), ns, d->hsps.data(), n))
            return HB_ERR_STATE;
        return HB_OK;
    }
    static void abandon(void *u) { self(u)->drop(); }
};
}  // namespace

int hb_cosearch(hb_ctx *c, const char *index_path, const uint8_t *seq, int32_t len,
                int32_t n_ctx, const int32_t *ctx_off, const int32_t *ctx_len,
                const int32_t *xdrop, const int32_t *cutoff, const hb_params *params,
                const uint8_t *lut_mask, const hb_runtime_options *options,
                int32_t cpu_threads, int32_t gapped, int64_t device_budget,
                hb_cpu_search_fn cpu, hb_gpu_results_fn gpu, void *user,
                hb_volume_report *reports, int32_t report_cap, int32_t *n_volumes)
{
    if (!c || !index_path || !cpu || !n_volumes) return HB_ERR_ARG;
    hb_query *q = nullptr;
    int rc = hb_query_prepare_masked(c, seq, len, n_ctx, ctx_off, ctx_len, xdrop, cutoff, params,
                                     lut_mask, &q);
    if (rc) return rc;
    CoschedDevice dev;
    dev.ctx = c;
    dev.q = q;
    dev.budget = device_budget;
    dev.gpu = gpu;
    dev.user = user;
    hb_cosched_steps steps;
    steps.user = &dev;
    steps.fits = &CoschedDevice::fits;
    steps.launch = &CoschedDevice::launch;
    steps.fetch = &CoschedDevice::fetch;
    steps.abandon = &CoschedDevice::abandon;
    rc = hb_cosched_run(index_path, options, cpu_threads, gapped, len,
                        (uint32_t)(len / std::max(n_ctx, 1)), &steps, cpu, user, reports,
                        report_cap, n_volumes);
    dev.drop();
    hb_query_free(q);
    return rc;
}

// ------------------------------------------------------------------ no-alloc scan
int hb_scan_noalloc(hb_ctx *c, hb_db *db, hb_query *q, hb_hsp *d_hsps, uint32_t hsp_cap,
                    uint32_t *d_counts, uint32_t *d_n_hsp, void *stream)
{
    if (!c || !db || !q || db->ctx != c || q->ctx != c || !d_n_hsp || (hsp_cap && !d_hsps))
        return HB_ERR_ARG;
    if (!c->scratch_dev) return HB_ERR_STATE;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
    const int place = state_place(q, db);
    if (place < 0) return HB_ERR_UNSUPPORTED;
    hb_launch_cfg cfg;
    int rc = hb_scan_config(c, false, q->gs.qlen_pad_max, q->gs.ctx_pad_max, q->gs.diag_len_max,
                            db->n_tiles, &cfg, place);
    if (rc) return rc;
    hb_scan_args a;
    rc = plain_scan_args(db, q, place, cfg, &a);
    if (rc) return rc;
    int blocks = cfg.blocks;
    if (place == HB_PLACE_GLOBAL) {
        // the state of every wavefront of the grid must fit the scratch:
        // fewer workgroups for long queries, none fits -> refused
        const size_t per_block = HB_GSTATE_WORD * (size_t)q->gs.diag_len_max * (size_t)cfg.waves;
        if (per_block > HB_NOALLOC_STATE_BYTES) return HB_ERR_UNSUPPORTED;
        blocks = std::min(blocks, (int)(HB_NOALLOC_STATE_BYTES / per_block));
        a.gstate = reinterpret_cast<uint32_t *>(c->scratch_dev + HB_NOALLOC_HEAD);
    }
    uint32_t *ctr = reinterpret_cast<uint32_t *>(c->scratch_dev);
    a.out_count = ctr;
    a.work_counter = ctr + 1;
    a.out = d_hsps;
    a.out_cap = hsp_cap;
    a.stats = d_counts;
    a.group_totals = nullptr;
    // the previous call's counters stay in the scratch: cleared by a kernel
    HB_CHECK(hb_launch_fill_u32(ctr, 0u, 2, st));
    if (db->n_tiles)
        HB_CHECK(hb_launch_scan(a, false, blocks, cfg.waves * HB_WAVE, cfg.lds, st));
    HB_CHECK(hipMemcpyAsync(d_n_hsp, ctr, sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    return HB_OK;
}

int hb_scan_free(hb_scan *s)
{
    if (!s) return HB_ERR_ARG;
    hipSetDevice(s->ctx->device);
    if (!s->waited) hipStreamSynchronize(s->ctx->stream);
    hb_pool_put(s->ctx, s->d_out, (size_t)s->cap * sizeof(hb_hsp));
    hb_pool_put(s->ctx, s->d_misc, 2 * sizeof(uint32_t));
    hb_pool_put(s->ctx, s->d_stats, (size_t)std::max<uint32_t>(s->db->n, 1) * 3 * sizeof(uint32_t));
    if (s->d_gstate) hb_pool_put(s->ctx, s->d_gstate, s->gstate_bytes);
    if (s->ev0) hipEventDestroy(s->ev0);
    if (s->ev1) hipEventDestroy(s->ev1);
    delete s;
    return HB_OK;
}

int hb_scan_state_mode(const hb_scan *s)
{
    if (!s) return HB_ERR_ARG;
    return s->place == HB_PLACE_GLOBAL ? HB_STATE_GLOBAL : HB_STATE_LDS;
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
    s->d_out = (hb_hsp *)hb_pool_get(c, (size_t)s->cap * sizeof(hb_hsp));
    s->d_misc = (uint32_t *)hb_pool_get(c, 2 * sizeof(uint32_t));
    s->d_stats = (uint32_t *)hb_pool_get(c, (size_t)std::max<uint32_t>(db->n, 1) * 3 * sizeof(uint32_t));
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
        hb_pool_put(c, s->d_out, (size_t)s->cap * sizeof(hb_hsp));
        s->cap = n + n / 8 + 1024;
        s->d_out = (hb_hsp *)hb_pool_get(c, (size_t)s->cap * sizeof(hb_hsp));
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
    const uint32_t ns = s->db->n;
    if (ns == 0) return HB_OK;
    hb_ctx *c = s->ctx;
    HB_CHECK(hipSetDevice(c->device));
    // the layout is built on the device (counts -> first indices, records
    // grouped by subject and ordered by (s_off, q_off)), then copied out
    constexpr uint32_t kBig = 4096;
    const uint32_t nb = (ns + 4095) / 4096;
    const size_t b_st = (size_t)ns * sizeof(hb_subject_stats);
    const size_t b_rec = (size_t)std::max<uint64_t>(n, 1) * sizeof(hb_hsp);
    const size_t b_scr = ((size_t)ns + nb + 1) * sizeof(uint32_t);
    const size_t b_big = kBig * sizeof(hb_fin_big);
    auto *d_st = static_cast<hb_subject_stats *>(hb_pool_get(c, b_st));
    auto *d_rec = static_cast<hb_hsp *>(hb_pool_get(c, b_rec));
    auto *d_scr = static_cast<uint32_t *>(hb_pool_get(c, b_scr));
    auto *d_big = static_cast<hb_fin_big *>(hb_pool_get(c, b_big));
    auto release = [&]() {
        hb_pool_put(c, d_st, b_st);
        hb_pool_put(c, d_rec, b_rec);
        hb_pool_put(c, d_scr, b_scr);
        hb_pool_put(c, d_big, b_big);
    };
    if (!d_st || !d_rec || !d_scr || !d_big) {
        release();
        return HB_ERR_NOMEM;
    }
    uint32_t *d_nbig = d_scr + ns + nb;
    hipStream_t st = c->stream;
    uint32_t nbig = 0;
    hipError_t e = hb_launch_fill_u32(d_nbig, 0u, 1, st);
    if (e == hipSuccess)
        e = hb_launch_finalize(s->d_stats, ns, s->d_out, (uint32_t)n, d_st, d_rec, d_scr, d_big,
                               kBig, d_nbig, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&nbig, d_nbig, sizeof(nbig), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && stats && hb_copy_out(c, stats, d_st, b_st, st) != HB_OK)
        e = hipErrorUnknown;
    if (e == hipSuccess && hsps && n &&
        hb_copy_out(c, hsps, d_rec, (size_t)n * sizeof(hb_hsp), st) != HB_OK)
        e = hipErrorUnknown;
    // subjects with many records are ordered here
    std::vector<hb_fin_big> big;
    if (e == hipSuccess && hsps && nbig) {
        if (nbig <= kBig) {
            big.resize(nbig);
            e = hipMemcpy(big.data(), d_big, nbig * sizeof(hb_fin_big), hipMemcpyDeviceToHost);
        } else {
            std::vector<hb_subject_stats> all;
            const hb_subject_stats *sp = stats;
            if (!sp) {
                all.resize(ns);
                e = hipMemcpy(all.data(), d_st, b_st, hipMemcpyDeviceToHost);
                sp = all.data();
            }
            for (uint32_t i = 0; e == hipSuccess && i < ns; ++i)
                if (sp[i].n_hsp > HB_FIN_SORT_MAX)
                    big.push_back(hb_fin_big{i, sp[i].n_hsp, (uint64_t)sp[i].first});
        }
    }
    release();
    if (e != hipSuccess) {
        hb_note_hip_error(e, "hb_scan_results", __LINE__);
        return HB_ERR_DEVICE;
    }
    for (const hb_fin_big &b : big)
        std::sort(hsps + b.first, hsps + b.first + b.n, [](const hb_hsp &x, const hb_hsp &y) {
            if (x.s_off != y.s_off) return x.s_off < y.s_off;
            return x.q_off < y.q_off;
        });
    return HB_OK;
}

int hb_scan_results_compact(hb_scan *s, hb_subject_hits *subj, uint64_t subj_cap, hb_hsp *hsps,
                            uint64_t hsp_cap, hb_scan_totals *totals)
{
    if (!s || !totals) return HB_ERR_ARG;
    int rc = hb_scan_wait(s);
    if (rc) return rc;
    std::memset(totals, 0, sizeof(*totals));
    const uint64_t n = s->n_out;
    totals->n_hsp = n;
    const uint32_t ns = s->db->n;
    if (ns == 0) return HB_OK;
    hb_ctx *c = s->ctx;
    HB_CHECK(hipSetDevice(c->device));
    // the same device layout as hb_scan_results, with the per-subject
    // entries compacted to the subjects with HSPs
    constexpr uint32_t kBig = 4096;
    const uint32_t nb = (ns + 4095) / 4096;
    const uint64_t max_subj = std::min<uint64_t>(ns, n);
    const size_t b_sub = (size_t)std::max<uint64_t>(max_subj, 1) * sizeof(hb_subject_hits);
    const size_t b_rec = (size_t)std::max<uint64_t>(n, 1) * sizeof(hb_hsp);
    const size_t b_scr = (2 * (size_t)nb + 1 + ns) * sizeof(uint32_t);
    const size_t b_big = kBig * sizeof(hb_fin_big);
    const size_t b_small = 32;   // two 64-bit totals, the big-subject count
    auto *d_sub = static_cast<hb_subject_hits *>(hb_pool_get(c, b_sub));
    auto *d_rec = static_cast<hb_hsp *>(hb_pool_get(c, b_rec));
    auto *d_scr = static_cast<uint32_t *>(hb_pool_get(c, b_scr));
    auto *d_big = static_cast<hb_fin_big *>(hb_pool_get(c, b_big));
    auto *d_small = static_cast<unsigned long long *>(hb_pool_get(c, b_small));
    auto release = [&]() {
        hb_pool_put(c, d_sub, b_sub);
        hb_pool_put(c, d_rec, b_rec);
        hb_pool_put(c, d_scr, b_scr);
        hb_pool_put(c, d_big, b_big);
        hb_pool_put(c, d_small, b_small);
    };
    if (!d_sub || !d_rec || !d_scr || !d_big || !d_small) {
        release();
        return HB_ERR_NOMEM;
    }
    uint32_t *d_nbig = reinterpret_cast<uint32_t *>(d_small + 2);
    hipStream_t st = c->stream;
    unsigned long long small[3] = {0, 0, 0};
    uint32_t n_subj = 0;
    hipError_t e = hb_launch_fill_u32(reinterpret_cast<uint32_t *>(d_small), 0u, 6, st);
    if (e == hipSuccess)
        e = hb_launch_finalize_compact(s->d_stats, ns, s->d_out, (uint32_t)n, d_sub, d_rec, d_scr,
                                       d_small, d_big, kBig, d_nbig, st);
    if (e == hipSuccess) e = hipMemcpyAsync(small, d_small, sizeof(small), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&n_subj, d_scr + 2 * nb, sizeof(n_subj), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    const uint32_t nbig = (uint32_t)small[2];
    if (e == hipSuccess) {
        totals->lookup_hits = small[0];
        totals->extensions = small[1];
        totals->n_subjects = n_subj;
        if ((subj && subj_cap < n_subj) || (hsps && hsp_cap < n)) {
            release();
            return HB_ERR_CAPACITY;
        }
    }
    std::vector<hb_subject_hits> host_sub;
    hb_subject_hits *sp = subj;
    if (e == hipSuccess && !sp && hsps && nbig > kBig) {   // the host needs the entries below
        host_sub.resize(n_subj);
        sp = host_sub.data();
    }
    if (e == hipSuccess && sp && n_subj &&
        hb_copy_out(c, sp, d_sub, (size_t)n_subj * sizeof(hb_subject_hits), st) != HB_OK)
        e = hipErrorUnknown;
    if (e == hipSuccess && hsps && n &&
        hb_copy_out(c, hsps, d_rec, (size_t)n * sizeof(hb_hsp), st) != HB_OK)
        e = hipErrorUnknown;
    // subjects with many records are ordered here
    std::vector<hb_fin_big> big;
    if (e == hipSuccess && hsps && nbig) {
        if (nbig <= kBig) {
            big.resize(nbig);
            e = hipMemcpy(big.data(), d_big, nbig * sizeof(hb_fin_big), hipMemcpyDeviceToHost);
        } else {
            for (uint32_t i = 0; i < n_subj; ++i)
                if (sp[i].n_hsp > HB_FIN_SORT_MAX)
                    big.push_back(hb_fin_big{sp[i].oid, sp[i].n_hsp, sp[i].first});
        }
    }
    release();
    if (e != hipSuccess) {
        hb_note_hip_error(e, "hb_scan_results_compact", __LINE__);
        return HB_ERR_DEVICE;
    }
    for (const hb_fin_big &b : big)
        std::sort(hsps + b.first, hsps + b.first + b.n, [](const hb_hsp &x, const hb_hsp &y) {
            if (x.s_off != y.s_off) return x.s_off < y.s_off;
            return x.q_off < y.q_off;
        });
    return HB_OK;
}

void hb_gap_params_default(hb_gap_params *p)
{
    if (!p) return;
    p->gap_open = 11;
    p->gap_extend = 1;
    p->x_dropoff = HB_DEFAULT_GAP_XDROP;
    p->reserved = 0;
}

static bool gap_params_ok(const hb_gap_params *gp)
{
    return gp && gp->gap_open >= 0 && gp->gap_extend >= 0 && gp->x_dropoff >= 0 &&
           gp->gap_open <= 1000 && gp->gap_extend <= 1000 && gp->x_dropoff <= 100000;
}

// The gapped stage on ungapped records already on the device: st3_in holds
// 3 counters per subject (the HSP count third), rec_in the n records in any
// order.  Lays the records out by subject, runs the main and wide passes,
// compacts and copies out.  Runs on the context's stream; the caller holds
// the context lock.
static int gapped_core(hb_ctx *c, const hb_db *db, const uint8_t *d_query, const int8_t *d_mat,
                       const std::vector<int32_t> &ctx_off, const std::vector<int32_t> &ctx_len,
                       const int32_t *gap_cutoff, const hb_gap_params *gp, const uint32_t *st3_in,
                       const hb_hsp *rec_in, uint64_t n, hb_gapped_stats *stats,
                       hb_gapped_hsp *hsps, uint64_t cap, uint64_t *n_hsp, float *ms_out,
                       uint32_t *wide_out)
{
    const uint32_t ns = db->n;
    const int32_t n_ctx = (int32_t)ctx_off.size();
    // pool buffers, all returned at the end
    std::vector<std::pair<void *, size_t>> held;
    bool oom = false;
    auto get = [&](size_t bytes) -> void * {
        bytes = std::max<size_t>(bytes, 64);
        void *p = hb_pool_get(c, bytes);
        if (!p) oom = true;
        else held.emplace_back(p, bytes);
        return p;
    };
    constexpr uint32_t kBig = 4096;
    const uint32_t nb = (ns + 4095) / 4096;
    const uint64_t nr = std::max<uint64_t>(n, 1);
    auto *d_ust = static_cast<hb_subject_stats *>(get((size_t)ns * sizeof(hb_subject_stats)));
    auto *d_urec = static_cast<hb_hsp *>(get(nr * sizeof(hb_hsp)));
    auto *d_scr = static_cast<uint32_t *>(get(((size_t)ns + nb + 1) * sizeof(uint32_t)));
    auto *d_big = static_cast<hb_fin_big *>(get(kBig * sizeof(hb_fin_big)));
    auto *d_st3 = static_cast<uint32_t *>(get((size_t)ns * 3 * sizeof(uint32_t)));
    auto *d_gout = static_cast<hb_gapped_hsp *>(get(nr * sizeof(hb_gapped_hsp)));
    auto *d_misc = static_cast<uint32_t *>(get(((size_t)ns + 2) * sizeof(uint32_t)));
    auto *d_ctx = static_cast<int32_t *>(get((size_t)3 * n_ctx * sizeof(int32_t)));
    auto *d_first = static_cast<uint32_t *>(get(((size_t)ns + nb + 1) * sizeof(uint32_t)));
    auto *d_gst = static_cast<hb_gapped_stats *>(get((size_t)ns * sizeof(hb_gapped_stats)));
    auto *d_gdense = static_cast<hb_gapped_hsp *>(get(nr * sizeof(hb_gapped_hsp)));
    auto *d_list = static_cast<uint32_t *>(get(((size_t)ns + 129) * sizeof(uint32_t)));
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (auto &h : held) hb_pool_put(c, h.first, h.second);
        for (hipEvent_t e : ev)
            if (e) hipEventDestroy(e);
    };
    if (oom) {
        release();
        return HB_ERR_NOMEM;
    }
    hipStream_t st = c->stream;
    std::vector<int32_t> hctx(3 * (size_t)n_ctx);
    for (int32_t i = 0; i < n_ctx; ++i) {
        hctx[i] = ctx_off[i];
        hctx[n_ctx + i] = ctx_len[i];
        hctx[2 * n_ctx + i] = gap_cutoff[i];
    }
    hb_gap_args a{};
    a.res = db->d_res;
    a.offs = db->d_offs;
    a.query = d_query;
    a.ctx = d_ctx;
    a.n_ctx = n_ctx;
    a.mat32 = d_mat;
    a.gap_open = gp->gap_open;
    a.gap_extend = gp->gap_extend;
    a.x_dropoff = gp->x_dropoff;
    a.ust = d_ust;
    a.urec = d_urec;
    a.n_subjects = ns;
    a.counter = d_misc;
    a.out = d_gout;
    a.st3 = d_st3;
    a.wide = d_misc + 1;
    a.list = d_list + 129;
    a.n_list = d_list + 128;
    uint32_t n_wide = 0;
    uint32_t *d_nbig = d_scr + ns + nb;
    hipError_t e = hb_launch_fill_u32(d_nbig, 0u, 1, st);
    if (e == hipSuccess)
        e = hb_launch_finalize(st3_in, ns, rec_in, (uint32_t)n, d_ust, d_urec, d_scr, d_big,
                               kBig, d_nbig, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_ctx, hctx.data(), hctx.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hb_launch_fill_u32(d_misc, 0u, 2, st);
    if (e == hipSuccess) e = hb_launch_gap_list(a, d_list + 129, d_list + 128, d_list, st);
    for (int i = 0; e == hipSuccess && i < 4; ++i) e = hipEventCreate(&ev[i]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], st);
    if (e == hipSuccess) e = hb_launch_gapped(a, false, c->n_cu * 8, st);
    if (e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (e == hipSuccess) e = hipMemcpyAsync(&n_wide, d_misc + 1, sizeof(n_wide),
                                            hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e == hipSuccess && n_wide) {
        // subjects that outgrew the LDS capacities: same code, global scratch
        std::vector<uint32_t> wl(n_wide);
        std::vector<hb_subject_stats> ust(ns);
        e = hipMemcpy(wl.data(), d_misc + 2, n_wide * sizeof(uint32_t), hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            e = hipMemcpy(ust.data(), d_ust, (size_t)ns * sizeof(hb_subject_stats),
                          hipMemcpyDeviceToHost);
        uint64_t longest = 0;
        uint32_t n_max = 1;
        for (uint32_t k : wl) {
            longest = std::max<uint64_t>(longest, db->h_offs[k + 1] - db->h_offs[k]);
            n_max = std::max(n_max, ust[k].n_hsp);
        }
        uint32_t ring = 64;
        while (ring < longest + 2) ring <<= 1;
        a.ring_wide = ring;
        a.n_max = n_max;
        a.scratch_words = (uint32_t)((2 * (uint64_t)ring + 7 * (uint64_t)n_max + 63) & ~63ull);
        const uint64_t per_wave = 4ull * a.scratch_words;
        uint64_t waves = std::min<uint64_t>(n_wide, 1024);
        waves = std::max<uint64_t>(1, std::min<uint64_t>(waves, ((uint64_t)1 << 30) / per_wave));
        const int blocks = (int)((waves + 3) / 4);
        a.scratch = static_cast<int32_t *>(get((size_t)blocks * 4 * per_wave));
        a.n_wide = n_wide;
        if (oom) e = hipErrorOutOfMemory;
        if (e == hipSuccess) e = hb_launch_fill_u32(d_misc, 0u, 1, st);
        if (e == hipSuccess) e = hipEventRecord(ev[2], st);
        if (e == hipSuccess) e = hb_launch_gapped(a, true, blocks, st);
        if (e == hipSuccess) e = hipEventRecord(ev[3], st);
    }
    if (wide_out) *wide_out = n_wide;
    if (e == hipSuccess) e = hb_launch_first_index(d_st3, ns, d_first, d_first + ns, st);
    if (e == hipSuccess) e = hb_launch_gap_compact(a, d_first, d_gst, d_gdense, st);
    uint32_t tail[2] = {0, 0};
    if (e == hipSuccess)
        e = hipMemcpyAsync(&tail[0], d_first + ns - 1, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(&tail[1], d_st3 + 3 * (size_t)(ns - 1) + 2, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float ms = 0.f, ms2 = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (e == hipSuccess && n_wide) e = hipEventElapsedTime(&ms2, ev[2], ev[3]);
    if (ms_out) *ms_out = ms + ms2;
    const uint64_t total = (uint64_t)tail[0] + tail[1];
    if (e == hipSuccess) {
        if (n_hsp) *n_hsp = total;
        if (hsps && cap < total) {
            release();
            return HB_ERR_CAPACITY;
        }
        if (stats && hb_copy_out(c, stats, d_gst, (size_t)ns * sizeof(hb_gapped_stats), st) !=
                         HB_OK)
            e = hipErrorUnknown;
        if (e == hipSuccess && hsps && total &&
            hb_copy_out(c, hsps, d_gdense, total * sizeof(hb_gapped_hsp), st) != HB_OK)
            e = hipErrorUnknown;
    }
    release();
    if (e != hipSuccess) {
        hb_note_hip_error(e, "gapped stage", __LINE__);
        return e == hipErrorOutOfMemory ? HB_ERR_NOMEM : HB_ERR_DEVICE;
    }
    return HB_OK;
}

int hb_scan_gapped(hb_scan *s, const hb_gap_params *gp, const int32_t *gap_cutoff,
                   hb_gapped_stats *stats, hb_gapped_hsp *hsps, uint64_t cap, uint64_t *n_hsp)
{
    if (!s || !gap_params_ok(gp) || !gap_cutoff) return HB_ERR_ARG;
    int rc = hb_scan_wait(s);
    if (rc) return rc;
    if (n_hsp) *n_hsp = 0;
    hb_ctx *c = s->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    s->gap_ms = 0.f;
    s->gap_wide = 0;
    if (s->db->n == 0) return HB_OK;
    HB_CHECK(hipSetDevice(c->device));
    return gapped_core(c, s->db, s->q->gs.desc[0].query, s->q->gs.d_mat, s->q->ctx_off,
                       s->q->ctx_len, gap_cutoff, gp, s->d_stats, s->d_out, s->n_out, stats, hsps,
                       cap, n_hsp, &s->gap_ms, &s->gap_wide);
}

int hb_gapped_search(hb_ctx *c, hb_db *db, const uint8_t *seq, int32_t len, int32_t n_ctx,
                     const int32_t *ctx_off, const int32_t *ctx_len, const int8_t *matrix,
                     const hb_hsp *in, uint64_t n_in, const hb_gap_params *gp,
                     const int32_t *gap_cutoff, hb_gapped_stats *stats, hb_gapped_hsp *hsps,
                     uint64_t cap, uint64_t *n_hsp)
{
    if (!c || !db || db->ctx != c || !gap_params_ok(gp) || !gap_cutoff || (n_in && !in))
        return HB_ERR_ARG;
    if (n_hsp) *n_hsp = 0;
    std::vector<int32_t> xd(n_ctx > 0 ? n_ctx : 1, 1);
    int rc = hb_check_query(seq, len, n_ctx, ctx_off, ctx_len, xd.data());
    if (rc) return rc;
    if (n_in >= 0xffffffffull) return HB_ERR_UNSUPPORTED;
    // every record must lie inside its context and subject, on its diagonal
    for (uint64_t i = 0; i < n_in; ++i) {
        const hb_hsp &h = in[i];
        if (h.oid >= db->n || h.context < 0 || h.context >= n_ctx || h.len < 1) return HB_ERR_ARG;
        const int64_t c0 = ctx_off[h.context], c1 = c0 + ctx_len[h.context];
        const int64_t slen = (int64_t)(db->h_offs[h.oid + 1] - db->h_offs[h.oid]);
        if (h.q_start < c0 || (int64_t)h.q_start + h.len > c1 || h.s_start < 0 ||
            (int64_t)h.s_start + h.len > slen || h.q_off - h.s_off != h.q_start - h.s_start)
            return HB_ERR_ARG;
    }
    hb_resolved_params rp;
    hb_params p;
    hb_params_default(&p);
    p.matrix = matrix;
    rc = hb_resolve_params(&p, &rp);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    if (db->n == 0) return HB_OK;
    const uint32_t ns = db->n;
    const size_t b_q = (size_t)len + 64, b_rec = std::max<size_t>(n_in, 1) * sizeof(hb_hsp);
    const size_t b_st3 = (size_t)ns * 3 * sizeof(uint32_t);
    auto *d_q = static_cast<uint8_t *>(hb_pool_get(c, b_q));
    auto *d_mat = static_cast<int8_t *>(hb_pool_get(c, 1024));
    auto *d_rec = static_cast<hb_hsp *>(hb_pool_get(c, b_rec));
    auto *d_st3 = static_cast<uint32_t *>(hb_pool_get(c, b_st3));
    auto release = [&]() {
        hb_pool_put(c, d_q, b_q);
        hb_pool_put(c, d_mat, 1024);
        hb_pool_put(c, d_rec, b_rec);
        hb_pool_put(c, d_st3, b_st3);
    };
    if (!d_q || !d_mat || !d_rec || !d_st3) {
        release();
        return HB_ERR_NOMEM;
    }
    hipStream_t st = c->stream;
    hipError_t e = hipMemcpyAsync(d_q, seq, (size_t)len, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_mat, rp.mat32, 1024, hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_in)
        e = hipMemcpyAsync(d_rec, in, n_in * sizeof(hb_hsp), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hb_launch_fill_u32(d_st3, 0u, 3 * ns, st);
    if (e == hipSuccess) e = hb_launch_count_oids(d_rec, (uint32_t)n_in, d_st3, st);
    if (e != hipSuccess) {
        release();
        hb_note_hip_error(e, "hb_gapped_search", __LINE__);
        return HB_ERR_DEVICE;
    }
    std::vector<int32_t> co(ctx_off, ctx_off + n_ctx), cl(ctx_len, ctx_len + n_ctx);
    rc = gapped_core(c, db, d_q, d_mat, co, cl, gap_cutoff, gp, d_st3, d_rec, n_in, stats, hsps,
                     cap, n_hsp, nullptr, nullptr);
    release();
    return rc;
}

int hb_scan_gapped_info(hb_scan *s, float *ms, uint32_t *n_wide)
{
    if (!s) return HB_ERR_ARG;
    if (ms) *ms = s->gap_ms;
    if (n_wide) *n_wide = s->gap_wide;
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

}  // extern "C"
