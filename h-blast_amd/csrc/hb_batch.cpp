// hb_batch.cpp -- batched form: many queries, one persistent launch per
// LDS size class, hipGraph capture/replay, autotuned group size.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

#include "hb_runtime.h"

namespace {

struct Class {
    int32_t g_begin = 0, g_end = 0;   // groups [g_begin, g_end) of the group set
    int32_t qlen_pad_max = 0, ctx_pad_max = 4, diag_len_max = 64;
    hb_launch_cfg cfg;
    uint32_t tiles_per_item = 1;
    int place = HB_PLACE_LDS;          // where the diagonal state goes
    uint32_t *d_gstate = nullptr;   // diag_len_max words per wavefront of the grid
    size_t gstate_bytes = 0;
};

}  // namespace

struct hb_batch {
    hb_ctx *ctx = nullptr;
    hb_db *db = nullptr;
    hipStream_t st = nullptr;                  // the context stream, or the caller's
    bool launched = false;                     // hb_batch_launch'ed, not yet finished
    bool reran = false;                        // the last finish reran an overflowed pass
    hb_resolved_params p;
    hb_batch_opts o;
    std::vector<hb_concat> queries;
    bool built = false;
    int32_t group_res = 0;
    float build_ms = 0;
    // launch geometry: 0 = the configuration's own choice
    int32_t g_waves = 0, g_ring = 0, g_items = 64;
    int32_t tune_candidates = 0;
    // groups
    hb_group_set gs;
    std::vector<Class> classes;
    std::vector<int32_t> ctx_query, ctx_local, q_base;   // global context -> query, local ctx
    // run state
    hb_hsp *d_out = nullptr;
    uint32_t cap = 0;
    uint32_t *d_misc = nullptr;                // [0] HSP count, [1 + c] work counter of class c
    unsigned long long *d_totals = nullptr;    // 3 per group
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float kernel_ms = 0;
    uint64_t tot[3] = {0, 0, 0};
    std::vector<hb_hsp> results;               // ordered by (query, oid, s_off, q_off)
    std::vector<uint64_t> q_first, q_count;
};

static void free_graph(hb_batch *b)
{
    if (b->gexec) hipGraphExecDestroy(b->gexec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

static void free_groups(hb_batch *b)
{
    free_graph(b);
    hb_group_set_free(&b->gs);
    hb_pool_put(b->ctx, b->d_misc, (1 + b->classes.size()) * sizeof(uint32_t));
    hb_pool_put(b->ctx, b->d_totals, 3 * b->gs.desc.size() * sizeof(unsigned long long));
    for (Class &c : b->classes)
        if (c.d_gstate) hb_pool_put(b->ctx, c.d_gstate, c.gstate_bytes);
    b->d_misc = nullptr;
    b->d_totals = nullptr;
    b->classes.clear();
    b->built = false;
}

// Packs queries into groups (first-fit decreasing by length; a query longer
// than the target forms its own group), concatenates each group with
// `window` code-0 residues between queries, and orders groups by diagonal
// array size so that groups of one LDS configuration are contiguous.
static void make_groups(hb_batch *b, int32_t target, std::vector<hb_concat> *groups,
                        std::vector<int32_t> *ctx_base, std::vector<int32_t> *diag)
{
    const int32_t pad = b->p.window;
    const int32_t nq = (int32_t)b->queries.size();
    std::vector<int32_t> ord(nq);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int32_t x, int32_t y) {
        return b->queries[x].seq.size() > b->queries[y].seq.size();
    });
    std::vector<std::vector<int32_t>> members;
    std::vector<int32_t> used;
    for (int32_t qi : ord) {
        const int32_t L = (int32_t)b->queries[qi].seq.size();
        int32_t slot = -1;
        for (size_t gi = 0; gi < members.size(); ++gi)
            if (used[gi] + pad + L <= target) {
                slot = (int32_t)gi;
                break;
            }
        if (slot < 0) {
            members.push_back({qi});
            used.push_back(L);
        } else {
            members[slot].push_back(qi);
            used[slot] += pad + L;
        }
    }
    // order groups by their diagonal array length (ascending)
    std::vector<int32_t> gorder(members.size());
    std::iota(gorder.begin(), gorder.end(), 0);
    std::stable_sort(gorder.begin(), gorder.end(),
                     [&](int32_t x, int32_t y) { return used[x] < used[y]; });
    b->q_base.assign(nq, 0);
    b->ctx_query.clear();
    b->ctx_local.clear();
    groups->clear();
    ctx_base->clear();
    diag->clear();
    for (int32_t gi : gorder) {
        hb_concat G;
        ctx_base->push_back((int32_t)b->ctx_query.size());
        bool masked = false;   // the group carries a mask if any member does
        for (int32_t qi : members[gi]) masked |= !b->queries[qi].mask.empty();
        for (size_t k = 0; k < members[gi].size(); ++k) {
            const int32_t qi = members[gi][k];
            const hb_concat &Q = b->queries[qi];
            if (k) G.seq.insert(G.seq.end(), (size_t)pad, (uint8_t)0);
            const int32_t base = (int32_t)G.seq.size();
            b->q_base[qi] = base;
            G.seq.insert(G.seq.end(), Q.seq.begin(), Q.seq.end());
            if (masked) {
                G.mask.resize((size_t)base, 0);
                if (Q.mask.empty()) G.mask.insert(G.mask.end(), Q.seq.size(), (uint8_t)0);
                else G.mask.insert(G.mask.end(), Q.mask.begin(), Q.mask.end());
            }
            for (size_t c = 0; c < Q.ctx_off.size(); ++c) {
                G.ctx_off.push_back(base + Q.ctx_off[c]);
                G.ctx_len.push_back(Q.ctx_len[c]);
                G.xdrop.push_back(Q.xdrop[c]);
                G.cutoff.push_back(Q.cutoff[c]);
                b->ctx_query.push_back(qi);
                b->ctx_local.push_back((int32_t)c);
            }
        }
        if (masked) G.mask.resize(G.seq.size(), 0);
        diag->push_back(((int32_t)G.seq.size() + b->p.window + 63) & ~63);
        groups->push_back(std::move(G));
    }
}

// Sizes work items for a pass over n_tiles tiles: about 64 items per
// workgroup for balance, few enough for the queue atomic.
static void size_items(hb_batch *b, uint32_t n_tiles)
{
    for (Class &c : b->classes) {
        const uint32_t n_groups = (uint32_t)(c.g_end - c.g_begin);
        const uint64_t want = (uint64_t)c.cfg.blocks * (uint64_t)std::max(b->g_items, 1);
        const uint64_t per_group = std::max<uint64_t>(1, want / std::max<uint32_t>(n_groups, 1));
        c.tiles_per_item = (uint32_t)std::max<uint64_t>(1, (n_tiles + per_group - 1) / per_group);
    }
}

// Wavefronts per CU that the scan reaches with one class's LDS layout
// (0 when no configuration fits).
static int class_waves(hb_batch *b, int32_t qlen_pad, int32_t ctx_pad, int32_t diag_len)
{
    const int place = hb_state_choose(b->ctx, diag_len, hb_scan_fixed_lds(qlen_pad, ctx_pad),
                                      b->db->max_len, b->p.window);
    hb_launch_cfg cfg;
    if (hb_scan_config(b->ctx, true, qlen_pad, ctx_pad, diag_len, 1u << 30, &cfg, place,
                       b->g_ring, b->g_waves) != HB_OK)
        return 0;
    return cfg.blocks / std::max(b->ctx->n_cu, 1) * cfg.waves;
}

// Splits the (diagonal-sorted) groups into classes, so that one launch's
// LDS layout, sized for its largest group, does not cost its other groups
// occupancy: a new class starts when the diagonal array would grow past
// twice the class's first one, or when a group would lower the wavefronts
// per CU the class reaches with its first group (packed groups sit just
// below the group size, and a class that also took the longest queries
// would run all of them at the occupancy of the longest).
static int build(hb_batch *b, int32_t target)
{
    free_groups(b);
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<hb_concat> groups;
    std::vector<int32_t> ctx_base, diag;
    make_groups(b, target, &groups, &ctx_base, &diag);
    // every query fits alone (checked when added); a group of several fits
    // unless the group size itself is too large for LDS
    for (const hb_concat &G : groups)
        if (!hb_query_lds_ok(b->ctx, (int32_t)G.seq.size(), (int32_t)G.ctx_off.size()))
            return HB_ERR_UNSUPPORTED;
    int rc = hb_group_set_build(b->ctx, groups, ctx_base, b->p, b->st, &b->gs);
    if (rc) return rc;
    const int32_t ng = (int32_t)groups.size();
    for (int32_t g = 0; g < ng;) {
        Class c;
        c.g_begin = g;
        const int32_t lim = std::max(diag[g] * 2, 1024);
        int w0 = -1;
        while (g < ng && diag[g] <= lim) {
            const int32_t qp = std::max(c.qlen_pad_max, b->gs.desc[g].qlen_pad);
            const int32_t cp = std::max(c.ctx_pad_max, hb_ctx_pad(b->gs.desc[g].n_ctx));
            const int32_t dl = std::max(c.diag_len_max, diag[g]);
            if (w0 < 0) {
                w0 = class_waves(b, qp, cp, dl);
            } else if (qp != c.qlen_pad_max || cp != c.ctx_pad_max || dl != c.diag_len_max) {
                if (class_waves(b, qp, cp, dl) < w0) break;
            }
            c.qlen_pad_max = qp;
            c.ctx_pad_max = cp;
            c.diag_len_max = dl;
            ++g;
        }
        c.g_end = g;
        b->classes.push_back(c);
    }
    for (Class &c : b->classes) {
        const uint32_t n_groups = (uint32_t)(c.g_end - c.g_begin);
        // long queries (a group of one query longer than the target): the
        // diagonal state goes to global memory as in the plain form
        c.place = hb_state_choose(b->ctx, c.diag_len_max,
                                  hb_scan_fixed_lds(c.qlen_pad_max, c.ctx_pad_max),
                                  b->db->max_len, b->p.window);
        rc = hb_scan_config(b->ctx, true, c.qlen_pad_max, c.ctx_pad_max, c.diag_len_max,
                            n_groups * b->db->n_tiles, &c.cfg, c.place, b->g_ring, b->g_waves);
        if (rc) return rc;
        if (c.place == HB_PLACE_GLOBAL) {
            c.gstate_bytes =
                (size_t)c.cfg.blocks * c.cfg.waves * (size_t)c.diag_len_max * HB_GSTATE_WORD;
            c.d_gstate = static_cast<uint32_t *>(hb_pool_get(b->ctx, c.gstate_bytes));
            if (!c.d_gstate) return HB_ERR_NOMEM;
        }
    }
    size_items(b, b->db->n_tiles);
    b->d_misc = static_cast<uint32_t *>(hb_pool_get(b->ctx, (1 + b->classes.size()) * sizeof(uint32_t)));
    b->d_totals = static_cast<unsigned long long *>(
        hb_pool_get(b->ctx, 3 * (size_t)ng * sizeof(unsigned long long)));
    if (!b->d_misc || !b->d_totals) return HB_ERR_NOMEM;
    b->group_res = target;
    b->built = true;
    b->build_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return HB_OK;
}

// Enqueues the whole scan (counter resets + one launch per class) on the
// batch's stream, over tiles [t_begin, t_end).
static int enqueue(hb_batch *b, uint32_t t_begin, uint32_t t_end)
{
    hb_db *db = b->db;
    const uint32_t nc = (uint32_t)b->classes.size();
    HB_CHECK(hb_launch_fill_u32(b->d_misc, 0u, 1 + nc, b->st));
    HB_CHECK(hb_launch_fill_u32(reinterpret_cast<uint32_t *>(b->d_totals), 0u,
                                (uint32_t)(6 * b->gs.desc.size()), b->st));
    for (uint32_t ci = 0; ci < nc; ++ci) {
        const Class &C = b->classes[ci];
        hb_scan_args a;
        memset(&a, 0, sizeof(a));
        a.res = db->d_scan_res;   // subjects in processing order (a.subj's offsets)
        a.offs = db->d_offs;
        a.subj = db->d_subj;
        a.tiles = db->d_tiles;
        a.tile_begin = t_begin;
        a.tile_end = t_end;
        a.tiles_per_item = C.tiles_per_item;
        a.groups = b->gs.d_desc + C.g_begin;
        a.n_groups = (uint32_t)(C.g_end - C.g_begin);
        a.qlen_pad_max = C.qlen_pad_max;
        a.ctx_pad_max = C.ctx_pad_max;
        a.diag_len_max = C.diag_len_max;
        a.wave_lds = (uint32_t)hb_scan_wave_lds(C.diag_len_max, C.cfg.ring,
                                                hb_state_bytes(C.place));
        a.ring = (uint32_t)C.cfg.ring;
        a.gstate = C.d_gstate;   // cleared by each wavefront when it binds a group
        a.state_limit = hb_state_limit(C.place, db->max_len);
        a.mat32 = b->gs.d_mat;
        a.window = b->p.window;
        a.work_counter = b->d_misc + 1 + ci;
        a.out = b->d_out;
        a.out_cap = b->cap;
        a.out_count = b->d_misc;
        a.stats = nullptr;
        a.group_totals = b->d_totals + 3 * (size_t)C.g_begin;
        if (t_end > t_begin)
            HB_CHECK(hb_launch_scan(a, true, C.cfg.blocks, C.cfg.waves * HB_WAVE, C.cfg.lds,
                                    b->st));
    }
    return HB_OK;
}

static int ensure_out(hb_batch *b, uint32_t cap)
{
    if (b->d_out && b->cap >= cap) return HB_OK;
    free_graph(b);   // the captured launches hold the old buffer
    hb_pool_put(b->ctx, b->d_out, (size_t)b->cap * sizeof(hb_hsp));
    b->cap = 0;
    b->d_out = static_cast<hb_hsp *>(hb_pool_get(b->ctx, (size_t)cap * sizeof(hb_hsp)));
    if (!b->d_out) return HB_ERR_NOMEM;
    b->cap = cap;
    return HB_OK;
}

// Enqueues one pass over tiles [t_begin, t_end) between the two timing
// events (replaying the captured graph for a full pass when enabled).
static int launch_pass(hb_batch *b, uint32_t t_begin, uint32_t t_end, bool use_graph)
{
    const bool full = t_begin == 0 && t_end == b->db->n_tiles;
    HB_CHECK(hipEventRecord(b->ev0, b->st));
    if (use_graph && full) {
        if (!b->gexec) {
            HB_CHECK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
            int rc = enqueue(b, t_begin, t_end);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(b->st, &g);
            if (rc) {
                if (g) hipGraphDestroy(g);
                return rc;
            }
            HB_CHECK(e);
            b->graph = g;
            HB_CHECK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        }
        HB_CHECK(hipGraphLaunch(b->gexec, b->st));
    } else {
        int rc = enqueue(b, t_begin, t_end);
        if (rc) return rc;
    }
    HB_CHECK(hipEventRecord(b->ev1, b->st));
    return HB_OK;
}

// Waits for the pass enqueued last; returns its HSP count, or grows the
// output buffer and returns *fits = false when the records did not fit
// (every HSP is counted, only cap are stored).
static int end_pass(hb_batch *b, uint32_t *n_out, bool *fits)
{
    uint32_t n = 0;
    HB_CHECK(hipMemcpyAsync(&n, b->d_misc, sizeof(n), hipMemcpyDeviceToHost, b->st));
    HB_CHECK(hipStreamSynchronize(b->st));
    *fits = n <= b->cap;
    if (*fits) {
        *n_out = n;
        HB_CHECK(hipEventElapsedTime(&b->kernel_ms, b->ev0, b->ev1));
        return HB_OK;
    }
    return ensure_out(b, n + n / 8 + 4096);
}

// One timed pass over tiles [t_begin, t_end); returns the HSP count.
static int run_pass(hb_batch *b, uint32_t t_begin, uint32_t t_end, bool use_graph, uint32_t *n_out)
{
    for (;;) {
        int rc = launch_pass(b, t_begin, t_end, use_graph);
        bool fits = false;
        if (!rc) rc = end_pass(b, n_out, &fits);
        if (rc || fits) return rc;
    }
}

// Times one candidate configuration on the slice [t0, t0 + span) of the
// database tiles (a warm pass, then a timed one); *ms = its kernel time, or
// a negative value when the configuration does not fit.
static int time_candidate(hb_batch *b, int32_t target, uint32_t t0, uint32_t span, float *ms)
{
    *ms = -1.f;
    int rc = build(b, target);
    if (rc == HB_ERR_UNSUPPORTED) return HB_OK;
    if (rc) return rc;
    // items sized for the slice, so that the timing is not a tail effect
    size_items(b, span);
    uint32_t n = 0;
    rc = run_pass(b, t0, t0 + span, false, &n);   // warm
    if (!rc) rc = run_pass(b, t0, t0 + span, false, &n);
    if (rc == HB_ERR_UNSUPPORTED) return HB_OK;
    if (rc) return rc;
    *ms = b->kernel_ms;
    b->tune_candidates++;
    return HB_OK;
}

// Autotuner: times candidates on a median-length slice of the database
// (4% of the tiles) and keeps the fastest, one knob at a time: the group
// size (query residues per group: more groups rescan the database, larger
// ones starve LDS), then wavefronts per workgroup, the staging ring (which
// trades LDS for the number of subjects whose extensions batch together)
// and the work items per workgroup (subjects per queue pop: balance against
// queue traffic).
static int autotune(hb_batch *b, int32_t *best)
{
    const uint32_t nt = b->db->n_tiles;
    uint32_t span = b->o.tune_tiles ? b->o.tune_tiles : std::max<uint32_t>(nt / 25, 64);
    span = std::min(span, nt);
    const uint32_t t0 = (nt - span) / 2;   // subjects of median length
    uint64_t total = 0;
    for (const hb_concat &q : b->queries) total += q.seq.size() + (uint64_t)b->p.window;
    float best_ms = 0, ms = 0;
    *best = 0;
    b->tune_candidates = 0;
    for (int32_t cand : {512, 768, 1024, 1536, 2048, 3072, 4096, 6144, 8192}) {
        int rc = time_candidate(b, cand, t0, span, &ms);
        if (rc) return rc;
        if (ms >= 0 && (*best == 0 || ms < best_ms)) {
            best_ms = ms;
            *best = cand;
        }
        if ((uint64_t)cand >= total) break;   // one group already holds every query
    }
    if (!*best) return HB_ERR_UNSUPPORTED;
    // launch geometry at the chosen group size, one knob at a time
    struct Knob {
        int32_t *field;
        std::vector<int32_t> values;
        bool user;
    } knobs[] = {{&b->g_waves, {2, 1}, b->o.waves_per_workgroup > 0},
                 {&b->g_ring, {1024, 1536}, b->o.ring_bytes > 0},
                 {&b->g_items, {16, 256}, b->o.items_per_workgroup > 0}};
    for (Knob &k : knobs) {
        if (k.user) continue;
        const int32_t keep = *k.field;
        int32_t pick = keep;
        for (int32_t v : k.values) {
            *k.field = v;
            int rc = time_candidate(b, *best, t0, span, &ms);
            if (rc) return rc;
            if (ms >= 0 && ms < best_ms * 0.98f) {   // a clear gain only
                best_ms = ms;
                pick = v;
            }
        }
        *k.field = pick;
    }
    return HB_OK;
}

// Fetches the n records and per-group counters of the last pass, maps
// records back to their queries (local context and offsets) and orders each
// query's records by (oid, s_off, q_off).
static int collect(hb_batch *b, uint32_t n)
{
    b->results.resize(n);
    std::vector<unsigned long long> tot(3 * b->gs.desc.size());
    HB_CHECK(hipMemcpyAsync(tot.data(), b->d_totals, tot.size() * sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, b->st));
    HB_CHECK(hipStreamSynchronize(b->st));
    if (n) {
        const int rc = hb_copy_out(b->ctx, b->results.data(), b->d_out, (size_t)n * sizeof(hb_hsp), b->st);
        if (rc) return rc;
    }
    b->tot[0] = b->tot[1] = b->tot[2] = 0;
    for (size_t g = 0; g < b->gs.desc.size(); ++g)
        for (int k = 0; k < 3; ++k) b->tot[k] += tot[3 * g + k];
    // partition by query (counting), then order each query's records on
    // host threads
    const size_t nq = b->queries.size();
    std::vector<int32_t> qid(n);
    b->q_first.assign(nq + 1, 0);
    b->q_count.assign(nq, 0);
    for (uint32_t i = 0; i < n; ++i) {
        hb_hsp &h = b->results[i];
        const int32_t qi = b->ctx_query[h.context];
        qid[i] = qi;
        h.context = b->ctx_local[h.context];
        h.q_start -= b->q_base[qi];
        h.q_off -= b->q_base[qi];
        b->q_count[qi]++;
    }
    for (size_t q = 0; q < nq; ++q) b->q_first[q + 1] = b->q_first[q] + b->q_count[q];
    std::vector<hb_hsp> sorted(n);
    {
        std::vector<uint64_t> cur(b->q_first.begin(), b->q_first.end() - 1);
        for (uint32_t i = 0; i < n; ++i) sorted[cur[qid[i]]++] = b->results[i];
    }
    const unsigned nt = std::max(1u, std::min<unsigned>({16u, std::thread::hardware_concurrency(),
                                                         (unsigned)std::max<size_t>(1, nq)}));
    auto order = [&](size_t t) {
        for (size_t q = t; q < nq; q += nt)
            std::sort(sorted.begin() + (long)b->q_first[q], sorted.begin() + (long)b->q_first[q + 1],
                      [](const hb_hsp &x, const hb_hsp &y) {
                          if (x.oid != y.oid) return x.oid < y.oid;
                          if (x.s_off != y.s_off) return x.s_off < y.s_off;
                          return x.q_off < y.q_off;
                      });
    };
    if (nt == 1 || n < 65536) {
        for (size_t t = 0; t < nt; ++t) order(t);
    } else {
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) th.emplace_back(order, (size_t)t);
        for (auto &x : th) x.join();
    }
    b->results.swap(sorted);
    return HB_OK;
}

int hb_batch_create_on(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_batch_opts *opts,
                       hipStream_t st, hb_batch **out)
{
    if (!ctx || !db || !out || db->ctx != ctx) return HB_ERR_ARG;
    *out = nullptr;
    hb_batch *b = new (std::nothrow) hb_batch;
    if (!b) return HB_ERR_NOMEM;
    b->ctx = ctx;
    b->db = db;
    hb_batch_opts_default(&b->o);
    if (opts) b->o = *opts;
    const hb_batch_opts &o = b->o;
    if ((o.waves_per_workgroup != 0 && o.waves_per_workgroup != 1 &&
         o.waves_per_workgroup != 2 && o.waves_per_workgroup != 4) ||
        (o.ring_bytes != 0 && (o.ring_bytes < 256 || o.ring_bytes % 64)) ||
        o.items_per_workgroup < 0 || o.group_residues < 0) {
        delete b;
        return HB_ERR_ARG;
    }
    b->g_waves = o.waves_per_workgroup;
    b->g_ring = o.ring_bytes;
    b->g_items = o.items_per_workgroup > 0 ? o.items_per_workgroup : 64;
    int rc = hb_resolve_params(params, &b->p);
    if (rc) {
        delete b;
        return rc;
    }
    std::lock_guard<std::mutex> lk(ctx->mu);
    b->st = st ? st : ctx->stream;
    if (hipSetDevice(ctx->device) != hipSuccess || hipEventCreate(&b->ev0) != hipSuccess ||
        hipEventCreate(&b->ev1) != hipSuccess) {
        if (b->ev0) hipEventDestroy(b->ev0);
        delete b;
        return HB_ERR_DEVICE;
    }
    *out = b;
    return HB_OK;
}

// Builds the groups (fixed group size: no autotuning on this path) and
// enqueues one pass on the batch's stream without waiting for it.
int hb_batch_launch(hb_batch *b)
{
    if (!b) return HB_ERR_ARG;
    if (b->queries.empty() || b->launched) return HB_ERR_STATE;
    hb_ctx *c = b->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    int rc = HB_OK;
    if (!b->built) {
        rc = build(b, b->o.group_residues > 0 ? b->o.group_residues : 768);
        if (rc) return rc;
    }
    rc = ensure_out(b, std::max<uint32_t>(b->cap, 4096));
    if (rc) return rc;
    rc = launch_pass(b, 0, b->db->n_tiles, false);
    if (rc) return rc;
    b->launched = true;
    return HB_OK;
}

bool hb_batch_ready(hb_batch *b)
{
    return !b->launched || hipEventQuery(b->ev1) != hipErrorNotReady;
}

// Waits for the launched pass, reruns it once with a larger buffer when
// the records did not fit, and collects the results.
int hb_batch_finish(hb_batch *b)
{
    if (!b) return HB_ERR_ARG;
    if (!b->launched) return HB_ERR_STATE;
    b->launched = false;
    hb_ctx *c = b->ctx;
    HB_CHECK(hipSetDevice(c->device));
    uint32_t n = 0;
    bool fits = false;
    int rc = end_pass(b, &n, &fits);
    if (rc) return rc;
    b->reran = !fits;
    if (!fits) {
        std::lock_guard<std::mutex> lk(c->mu);
        rc = run_pass(b, 0, b->db->n_tiles, false, &n);
        if (rc) return rc;
    }
    return collect(b, n);
}

bool hb_batch_rerun_needed(const hb_batch *b) { return b->reran; }

size_t hb_batch_device_bytes(const hb_batch *b)
{
    return b->gs.fixed_bytes + b->gs.entries_bytes + (size_t)b->cap * sizeof(hb_hsp) +
           b->gs.desc.size() * (sizeof(hb_group_desc) + 3 * sizeof(unsigned long long));
}

// Sets the output capacity (records) before launch.
int hb_batch_reserve(hb_batch *b, uint32_t cap)
{
    if (!b || b->launched) return HB_ERR_STATE;
    std::lock_guard<std::mutex> lk(b->ctx->mu);
    HB_CHECK(hipSetDevice(b->ctx->device));
    return ensure_out(b, cap);
}

extern "C" {

void hb_batch_opts_default(hb_batch_opts *o)
{
    if (!o) return;
    o->group_residues = 0;
    o->use_graph = 1;
    o->tune_tiles = 0;
    o->waves_per_workgroup = 0;
    o->ring_bytes = 0;
    o->items_per_workgroup = 0;
}

int hb_batch_create(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_batch_opts *opts,
                    hb_batch **out)
{
    return hb_batch_create_on(ctx, db, params, opts, nullptr, out);
}

int hb_batch_add(hb_batch *b, const uint8_t *seq, int32_t len, int32_t n_ctx,
                 const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                 const int32_t *cutoff, int32_t *query_id)
{
    return hb_batch_add_masked(b, seq, len, n_ctx, ctx_off, ctx_len, xdrop, cutoff, nullptr,
                               query_id);
}

int hb_batch_add_masked(hb_batch *b, const uint8_t *seq, int32_t len, int32_t n_ctx,
                        const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                        const int32_t *cutoff, const uint8_t *lut_mask, int32_t *query_id)
{
    if (!b || !cutoff) return HB_ERR_ARG;
    if (b->built) return HB_ERR_STATE;
    int rc = hb_check_query(seq, len, n_ctx, ctx_off, ctx_len, xdrop);
    if (rc) return rc;
    // the query (and its group) must fit a workgroup's LDS: refused here, not
    // when the batch is built
    if (!hb_query_lds_ok(b->ctx, len, n_ctx)) return HB_ERR_UNSUPPORTED;
    hb_concat q;
    q.seq.assign(seq, seq + len);
    q.ctx_off.assign(ctx_off, ctx_off + n_ctx);
    q.ctx_len.assign(ctx_len, ctx_len + n_ctx);
    q.xdrop.assign(xdrop, xdrop + n_ctx);
    q.cutoff.assign(cutoff, cutoff + n_ctx);
    if (lut_mask) q.mask.assign(lut_mask, lut_mask + len);
    b->queries.push_back(std::move(q));
    if (query_id) *query_id = (int32_t)b->queries.size() - 1;
    return HB_OK;
}

int hb_batch_run(hb_batch *b)
{
    if (!b) return HB_ERR_ARG;
    if (b->queries.empty() || b->launched) return HB_ERR_STATE;
    hb_ctx *c = b->ctx;
    std::lock_guard<std::mutex> lk(c->mu);
    HB_CHECK(hipSetDevice(c->device));
    int rc = HB_OK;
    if (!b->built) {
        int32_t target = b->o.group_residues;
        if (target <= 0) {
            rc = ensure_out(b, 1u << 20);
            if (!rc) rc = autotune(b, &target);
            if (rc) return rc;
        }
        rc = build(b, target);
        if (rc) return rc;
    }
    rc = ensure_out(b, std::max<uint32_t>(b->cap, std::max<uint32_t>(1u << 20, b->db->n / 2)));
    if (rc) return rc;
    uint32_t n = 0;
    rc = run_pass(b, 0, b->db->n_tiles, b->o.use_graph != 0, &n);
    if (rc) return rc;
    return collect(b, n);
}

int hb_batch_results(hb_batch *b, int32_t query_id, hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp)
{
    if (!b || query_id < 0 || (size_t)query_id >= b->queries.size()) return HB_ERR_ARG;
    if (b->q_count.size() != b->queries.size()) return HB_ERR_STATE;
    const uint64_t n = b->q_count[query_id];
    if (n_hsp) *n_hsp = n;
    if (!hsps) return HB_OK;
    if (cap < n) return HB_ERR_CAPACITY;
    if (n) memcpy(hsps, &b->results[b->q_first[query_id]], (size_t)n * sizeof(hb_hsp));
    return HB_OK;
}

int hb_batch_info_get(hb_batch *b, hb_batch_info *info)
{
    if (!b || !info) return HB_ERR_ARG;
    info->n_queries = (int32_t)b->queries.size();
    info->n_groups = (int32_t)b->gs.desc.size();
    info->n_classes = (int32_t)b->classes.size();
    info->group_residues = b->group_res;
    info->kernel_ms = b->kernel_ms;
    info->build_ms = b->build_ms;
    info->lookup_hits = b->tot[0];
    info->extensions = b->tot[1];
    info->hsps = b->tot[2];
    info->waves_per_workgroup = 0;
    info->ring_bytes = 0;
    for (const Class &c : b->classes)   // the largest class's configuration
        if (c.g_end - c.g_begin >= 0 && c.cfg.lds) {
            info->waves_per_workgroup = c.cfg.waves;
            info->ring_bytes = c.cfg.ring;
        }
    info->items_per_workgroup = b->g_items;
    info->tune_candidates = b->tune_candidates;
    return HB_OK;
}

int hb_batch_free(hb_batch *b)
{
    if (!b) return HB_ERR_ARG;
    hipSetDevice(b->ctx->device);
    hipStreamSynchronize(b->st);
    free_groups(b);
    hb_pool_put(b->ctx, b->d_out, (size_t)b->cap * sizeof(hb_hsp));
    if (b->ev0) hipEventDestroy(b->ev0);
    if (b->ev1) hipEventDestroy(b->ev1);
    delete b;
    return HB_OK;
}

}  // extern "C"
