// hb_serve.cpp -- served form: a request loop over the resident database.
//
// Requests are queued by hb_serve_submit (any thread).  One server thread
// owns the device side: it admits queued requests into a batch while the
// batch stays within max_batch_queries / max_batch_residues and the device
// bytes of everything admitted stay within the memory budget, launches the
// batch on one of its streams without waiting, and polls the batches in
// flight.  A finished batch's results are split per request and the waiters
// are woken.  With max_inflight > 1 the next batch is launched while the
// previous one still runs, so the device starts it as workgroups of the
// earlier launch retire (continuous batching); no request waits for a
// batch boundary it did not need.
//
// A query's HSPs do not depend on its batch: queries are separated by
// `window` code-0 residues inside a group (no two-hit pairing and no
// extension crosses them) and each query's records are reported in
// (oid, s_off, q_off) order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hb_runtime.h"

// Completed requests whose latencies the percentiles cover (the most recent).
static constexpr size_t kLatencyWindow = 65536;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    uint64_t ticket = 0;
    hb_concat q;
    int64_t residues = 0;
    uint64_t est_hsps = 0;
    size_t est_bytes = 0;
    Clock::time_point t_submit;
    int state = 0;   // 0 queued, 1 on the device, 2 done
    int rc = HB_OK;
    std::vector<hb_hsp> hsps;
};

struct InFlight {
    hb_batch *b = nullptr;
    int slot = 0;   // stream index
    size_t bytes = 0;
    std::vector<std::shared_ptr<Request>> reqs;
};

}  // namespace

struct hb_server {
    hb_ctx *ctx = nullptr;
    hb_db *db = nullptr;
    hb_params params;
    hb_serve_opts o;
    int32_t window = 40;
    std::vector<hipStream_t> streams;
    std::vector<bool> slot_busy;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> live;
    std::vector<InFlight> inflight;   // touched by the server thread only
    size_t bytes_admitted = 0;
    bool stopping = false;
    uint64_t next_ticket = 1;

    // statistics (under mu)
    // latencies of the most recent completed requests (a ring of
    // kLatencyWindow), for the percentiles; count, sum and max cover all
    std::vector<double> lat_ms;
    size_t lat_next = 0;
    uint64_t lat_count = 0;
    double lat_sum = 0, lat_max = 0;
    hb_serve_stats st;
    Clock::time_point t_start;

    std::thread th;
};

// Device bytes one request adds to a batch: its output reservation, the
// tables of one group (a query is never split; it shares a group or opens
// one of its own), its lookup-table entries (about 30 per residue for
// BLOSUM62 at T 11; 64 are reserved) and, for a query long enough that its
// group keeps the diagonal state in global memory, that state.
static void estimate(const hb_server *s, Request *r)
{
    const double cells = (double)r->residues * (double)s->db->total;
    r->est_hsps = (uint64_t)(s->o.hsp_per_cell * cells) + 256;
    const size_t group_fixed = HB_NCELLS * (sizeof(uint2) + sizeof(uint32_t)) + 8192 +
                               HB_PV_WORDS * sizeof(uint32_t);
    r->est_bytes = r->est_hsps * sizeof(hb_hsp) + group_fixed +
                   64 * sizeof(uint32_t) * (size_t)r->residues;
    const int32_t D = (int32_t)((r->residues + s->window + 63) & ~63);
    if (hb_state_choose(s->ctx, D, hb_scan_fixed_lds((int32_t)((r->residues + 15) & ~15), 8),
                        s->db->max_len, s->window) == HB_PLACE_GLOBAL)
        r->est_bytes += hb_gstate_bytes_max(s->ctx, D);
}

static void finish_requests(hb_server *s, InFlight &f, int rc)
{
    const auto now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (size_t i = 0; i < f.reqs.size(); ++i) {
        Request &r = *f.reqs[i];
        r.rc = rc;
        if (rc == HB_OK) {
            uint64_t n = 0;
            hb_batch_results(f.b, (int32_t)i, nullptr, 0, &n);
            r.hsps.resize(n);
            if (n) r.rc = hb_batch_results(f.b, (int32_t)i, r.hsps.data(), n, &n);
        }
        r.state = 2;
        const double ms = std::chrono::duration<double, std::milli>(now - r.t_submit).count();
        if (r.rc == HB_OK) {
            if (s->lat_ms.size() < kLatencyWindow) {
                s->lat_ms.push_back(ms);
            } else {
                s->lat_ms[s->lat_next] = ms;
                s->lat_next = (s->lat_next + 1) % kLatencyWindow;
            }
            s->lat_count++;
            s->lat_sum += ms;
            s->lat_max = std::max(s->lat_max, ms);
            s->st.completed++;
            s->st.query_residues += (uint64_t)r.residues;
        }
    }
    if (rc == HB_OK) {
        hb_batch_info info;
        hb_batch_info_get(f.b, &info);
        s->st.gpu_ms += info.kernel_ms;
        if (hb_batch_rerun_needed(f.b)) s->st.reruns++;
    }
    s->bytes_admitted -= f.bytes;
    s->slot_busy[f.slot] = false;
    s->cv_done.notify_all();
}

// Forms and launches one batch from the head of the queue, if the budget
// allows; returns false when nothing was launched.
static bool admit(hb_server *s)
{
    InFlight f;
    uint64_t hsps = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->pending.empty()) return false;
        int slot = -1;
        for (size_t k = 0; k < s->slot_busy.size(); ++k)
            if (!s->slot_busy[k]) {
                slot = (int)k;
                break;
            }
        if (slot < 0) return false;
        // a batch that is not full waits up to linger_us for more requests
        if (!s->stopping && s->o.linger_us > 0 &&
            (int32_t)s->pending.size() < s->o.max_batch_queries) {
            int64_t queued = 0;
            for (const auto &r : s->pending) queued += r->residues;
            const auto waited = std::chrono::duration_cast<std::chrono::microseconds>(
                Clock::now() - s->pending.front()->t_submit).count();
            if (queued < s->o.max_batch_residues && waited < s->o.linger_us) return false;
        }
        int64_t res = 0;
        while (!s->pending.empty() && (int32_t)f.reqs.size() < s->o.max_batch_queries) {
            Request &r = *s->pending.front();
            if (!f.reqs.empty() && res + r.residues > s->o.max_batch_residues) break;
            if (s->bytes_admitted + f.bytes + r.est_bytes > (size_t)s->o.mem_budget) break;
            res += r.residues;
            f.bytes += r.est_bytes;
            hsps += r.est_hsps;
            r.state = 1;
            f.reqs.push_back(s->pending.front());
            s->pending.pop_front();
        }
        if (f.reqs.empty()) return false;   // waits for memory held by running batches
        f.slot = slot;
        s->slot_busy[slot] = true;
        s->bytes_admitted += f.bytes;
        s->st.peak_bytes = std::max<uint64_t>(s->st.peak_bytes, s->bytes_admitted);
        s->st.batches++;
    }
    hb_batch_opts bo;
    hb_batch_opts_default(&bo);
    bo.group_residues = s->o.group_residues;
    bo.use_graph = 0;   // every batch is launched once
    int rc = hb_batch_create_on(s->ctx, s->db, &s->params, &bo, s->streams[f.slot], &f.b);
    for (size_t i = 0; rc == HB_OK && i < f.reqs.size(); ++i) {
        const hb_concat &q = f.reqs[i]->q;
        rc = hb_batch_add_masked(f.b, q.seq.data(), (int32_t)q.seq.size(),
                                 (int32_t)q.ctx_off.size(), q.ctx_off.data(), q.ctx_len.data(),
                                 q.xdrop.data(), q.cutoff.data(),
                                 q.mask.empty() ? nullptr : q.mask.data(), nullptr);
    }
    if (rc == HB_OK)
        rc = hb_batch_reserve(f.b, (uint32_t)std::min<uint64_t>(hsps, 0xF0000000ull));
    if (rc == HB_OK) rc = hb_batch_launch(f.b);
    if (rc != HB_OK) {
        finish_requests(s, f, rc);
        if (f.b) hb_batch_free(f.b);
        return true;
    }
    s->inflight.push_back(std::move(f));
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.peak_inflight = std::max<int32_t>(s->st.peak_inflight, (int32_t)s->inflight.size());
    return true;
}

static void loop(hb_server *s)
{
    hipSetDevice(s->ctx->device);
    for (;;) {
        // collect finished batches, oldest first
        for (size_t i = 0; i < s->inflight.size();) {
            InFlight &f = s->inflight[i];
            if (!hb_batch_ready(f.b)) {
                ++i;
                continue;
            }
            const int rc = hb_batch_finish(f.b);
            finish_requests(s, f, rc);
            hb_batch_free(f.b);
            s->inflight.erase(s->inflight.begin() + (long)i);
        }
        while (admit(s)) {
        }
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->stopping && s->pending.empty() && s->inflight.empty()) break;
        if (s->inflight.empty() && s->pending.empty()) {
            // idle: sleep until a submit or stop
            s->cv_work.wait(lk, [&] { return s->stopping || !s->pending.empty(); });
        } else {
            // poll the batches in flight / the lingering batch
            s->cv_work.wait_for(lk, std::chrono::microseconds(200));
        }
    }
}

extern "C" {

void hb_serve_opts_default(hb_serve_opts *o)
{
    if (!o) return;
    o->max_batch_queries = 64;
    o->max_inflight = 2;
    o->max_batch_residues = 16384;
    o->mem_budget = 0;
    o->group_residues = 768;
    o->max_pending = 65536;
    o->hsp_per_cell = 5e-7;
    o->linger_us = 0;
}

int hb_serve_start(hb_ctx *ctx, hb_db *db, const hb_params *params, const hb_serve_opts *opts,
                   hb_server **out)
{
    if (!ctx || !db || !out || db->ctx != ctx) return HB_ERR_ARG;
    *out = nullptr;
    hb_serve_opts o;
    hb_serve_opts_default(&o);
    if (opts) o = *opts;
    if (o.max_batch_queries < 1 || o.max_inflight < 1 || o.max_inflight > 8 ||
        o.max_batch_residues < 1 || o.mem_budget < 0 || o.group_residues < 0 ||
        o.max_pending < 1 || !(o.hsp_per_cell >= 0) || o.linger_us < 0)
        return HB_ERR_ARG;
    hb_resolved_params rp;
    int rc = hb_resolve_params(params, &rp);
    if (rc) return rc;
    if (o.group_residues == 0) o.group_residues = 768;
    HB_CHECK(hipSetDevice(ctx->device));
    if (o.mem_budget == 0) {
        size_t fr = 0, tot = 0;
        HB_CHECK(hipMemGetInfo(&fr, &tot));
        o.mem_budget = (int64_t)(fr / 2);
    }
    hb_server *s = new (std::nothrow) hb_server;
    if (!s) return HB_ERR_NOMEM;
    s->ctx = ctx;
    s->db = db;
    hb_params_default(&s->params);
    if (params) s->params = *params;
    s->o = o;
    s->window = rp.window;
    memset(&s->st, 0, sizeof(s->st));
    for (int k = 0; k < o.max_inflight; ++k) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            for (hipStream_t x : s->streams) hipStreamDestroy(x);
            delete s;
            return HB_ERR_DEVICE;
        }
        s->streams.push_back(st);
    }
    s->slot_busy.assign(o.max_inflight, false);
    s->t_start = Clock::now();
    s->th = std::thread(loop, s);
    *out = s;
    return HB_OK;
}

int hb_serve_submit(hb_server *s, const uint8_t *seq, int32_t len, int32_t n_ctx,
                    const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                    const int32_t *cutoff, uint64_t *ticket)
{
    return hb_serve_submit_masked(s, seq, len, n_ctx, ctx_off, ctx_len, xdrop, cutoff, nullptr,
                                  ticket);
}

int hb_serve_submit_masked(hb_server *s, const uint8_t *seq, int32_t len, int32_t n_ctx,
                           const int32_t *ctx_off, const int32_t *ctx_len, const int32_t *xdrop,
                           const int32_t *cutoff, const uint8_t *lut_mask, uint64_t *ticket)
{
    if (!s || !cutoff || !ticket) return HB_ERR_ARG;
    int rc = hb_check_query(seq, len, n_ctx, ctx_off, ctx_len, xdrop);
    if (rc) return rc;
    // a query that cannot fit a workgroup's LDS is refused here, so it never
    // fails the batch it would have joined
    if (!hb_query_lds_ok(s->ctx, len, n_ctx)) return HB_ERR_UNSUPPORTED;
    auto r = std::make_shared<Request>();
    r->q.seq.assign(seq, seq + len);
    if (lut_mask) r->q.mask.assign(lut_mask, lut_mask + len);
    r->q.ctx_off.assign(ctx_off, ctx_off + n_ctx);
    r->q.ctx_len.assign(ctx_len, ctx_len + n_ctx);
    r->q.xdrop.assign(xdrop, xdrop + n_ctx);
    r->q.cutoff.assign(cutoff, cutoff + n_ctx);
    r->residues = len;
    estimate(s, r.get());
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return HB_ERR_STATE;
    s->st.submitted++;
    if (r->est_bytes > (size_t)s->o.mem_budget) {
        s->st.rejected++;
        return HB_ERR_CAPACITY;
    }
    if ((int64_t)s->pending.size() >= s->o.max_pending) {
        s->st.rejected++;
        return HB_ERR_BUSY;
    }
    r->ticket = s->next_ticket++;
    r->t_submit = Clock::now();
    *ticket = r->ticket;
    s->live.emplace(r->ticket, r);
    s->pending.push_back(std::move(r));
    s->cv_work.notify_one();
    return HB_OK;
}

int hb_serve_wait(hb_server *s, uint64_t ticket, int32_t timeout_ms)
{
    if (!s) return HB_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return HB_ERR_ARG;
    std::shared_ptr<Request> r = it->second;
    auto done = [&] { return r->state == 2; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
        return HB_ERR_BUSY;
    return r->rc;
}

int hb_serve_results(hb_server *s, uint64_t ticket, hb_hsp *hsps, uint64_t cap, uint64_t *n_hsp)
{
    if (!s) return HB_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(ticket);
    if (it == s->live.end()) return HB_ERR_ARG;
    Request &r = *it->second;
    if (r.state != 2) return HB_ERR_STATE;
    if (r.rc != HB_OK) {   // a failed request is released with its error
        const int rc = r.rc;
        s->live.erase(it);
        return rc;
    }
    if (n_hsp) *n_hsp = r.hsps.size();
    if (!hsps) return HB_OK;
    if (cap < r.hsps.size()) return HB_ERR_CAPACITY;
    if (!r.hsps.empty()) memcpy(hsps, r.hsps.data(), r.hsps.size() * sizeof(hb_hsp));
    s->live.erase(it);
    return HB_OK;
}

int hb_serve_stats_get(hb_server *s, hb_serve_stats *out)
{
    if (!s || !out) return HB_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - s->t_start).count();
    std::vector<double> v = s->lat_ms;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        // nearest-rank percentiles
        auto pct = [&](double p) {
            size_t k = (size_t)std::ceil(p / 100.0 * (double)v.size());
            return v[std::min(v.size(), std::max<size_t>(k, 1)) - 1];
        };
        out->p50_ms = pct(50);
        out->p99_ms = pct(99);
        out->mean_ms = s->lat_sum / (double)s->lat_count;
        out->max_ms = s->lat_max;
    }
    return HB_OK;
}

int hb_serve_stop(hb_server *s)
{
    if (!s) return HB_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv_work.notify_all();
    }
    if (s->th.joinable()) s->th.join();
    hipSetDevice(s->ctx->device);
    for (hipStream_t x : s->streams) hipStreamDestroy(x);
    delete s;
    return HB_OK;
}

}  // extern "C"
