// Sanitizer driver (test infrastructure): random and malformed inputs through
// the host-only C-ABI helpers (h-blast_amd/csrc/hb_seq.cpp, hb_hdb.cpp,
// hb_balance.cpp, hb_seg.cpp) and the CPU oracle (oracle/hbo.c, hbo_gap.c,
// hbo_hdb.c, hbo_balance.c, hbo_seg.c).  tests/test_sanitizers.py builds it
// with -fsanitize=address,undefined (and the threaded oracle parts with
// -fsanitize=thread) and fails on any report.  Exit status 0 = clean.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "hb_seed.h"
extern "C" {
#include "hbo.h"
}

static int fails = 0;
#define CHECK(c)                                                         \
    do {                                                                 \
        if (!(c)) {                                                      \
            std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                     \
        }                                                                \
    } while (0)

static void fasta(std::mt19937_64 &rng, int rounds)
{
    const char alphabet[] = "ACDEFGHIKLMNPQRSTVWYXBZUOJ*-acdefg \t\r0123456789>>\n\n\n;#.";
    for (int r = 0; r < rounds; ++r) {
        const size_t len = rng() % 600;
        std::string t;
        if (rng() % 2) t = ">";
        for (size_t i = 0; i < len; ++i) t += alphabet[rng() % (sizeof(alphabet) - 1)];
        uint64_t nr = 0;
        uint32_t ns = 0;
        const int rc = hb_fasta_parse(t.data(), t.size(), nullptr, 0, nullptr, nullptr, nullptr, 0,
                                      &nr, &ns);
        if (rc != HB_OK) {
            CHECK(rc == HB_ERR_RESIDUE && nr < t.size());
            continue;
        }
        for (int shrink = 0; shrink < 3; ++shrink) {
            const uint64_t rc_cap = nr > (uint64_t)shrink ? nr - shrink : 0;
            const uint32_t sc = ns > (uint32_t)shrink ? ns - shrink : 0;
            std::vector<uint8_t> res(rc_cap);
            std::vector<uint64_t> offs(sc + 1), tpos(sc);
            std::vector<uint32_t> tlen(sc);
            uint64_t nr2 = 0;
            uint32_t ns2 = 0;
            const int rc2 = hb_fasta_parse(t.data(), t.size(), res.data(), rc_cap, offs.data(),
                                           tpos.data(), tlen.data(), sc, &nr2, &ns2);
            CHECK(nr2 == nr && ns2 == ns);
            CHECK(shrink == 0 || nr == 0 ? rc2 == HB_OK || rc2 == HB_ERR_CAPACITY
                                         : rc2 == HB_ERR_CAPACITY);
            if (shrink == 0) {
                CHECK(rc2 == HB_OK);
                for (uint32_t k = 0; k < ns; ++k)
                    CHECK(offs[k] <= offs[k + 1] && tpos[k] + tlen[k] <= t.size());
                for (uint64_t k = 0; k < nr; ++k) CHECK(res[k] < 28);
            }
        }
    }
}

static void translate(std::mt19937_64 &rng, int rounds)
{
    const char nts[] = "ACGTacgtUNRY-";
    for (int r = 0; r < rounds; ++r) {
        const int32_t n = (int32_t)(rng() % 2000);
        std::string nt;
        for (int32_t i = 0; i < n; ++i) nt += nts[rng() % (sizeof(nts) - 1)];
        std::vector<uint8_t> out((size_t)hb_blastx_concat_len(n));
        int32_t off[6], len[6];
        CHECK(hb_blastx_translate(nt.data(), n, out.data(), off, len) == HB_OK);
        for (int f = 0; f < 6; ++f) CHECK(off[f] >= 0 && off[f] + len[f] <= (int32_t)out.size());
        std::string prot;
        for (int i = 0; i < 50; ++i) prot += "ACDXZ*-J#"[rng() % 9];
        std::vector<uint8_t> codes(prot.size());
        const int rc = hb_encode_protein(prot.data(), (int32_t)prot.size(), codes.data());
        CHECK(rc == HB_OK || rc == HB_ERR_RESIDUE);
    }
}

static void oracle(std::mt19937_64 &rng, int rounds, int threads)
{
    for (int r = 0; r < rounds; ++r) {
        const int32_t qlen = 20 + (int32_t)(rng() % 300);
        std::vector<uint8_t> q(qlen);
        for (auto &x : q) x = (uint8_t)(1 + rng() % 27);
        int32_t ctx_off[2] = {0, qlen / 2 + 1}, ctx_len[2] = {qlen / 2, qlen - qlen / 2 - 1};
        q[qlen / 2] = 0;
        int32_t xd[2] = {16, 16}, cut[2] = {30, 35};
        std::vector<uint8_t> mask(qlen, 0);
        for (int i = 0; i < 3; ++i) mask[rng() % qlen] = 1;
        hbo_query Q = {q.data(), qlen, 2, ctx_off, ctx_len, xd, cut, 3, 11, 40, nullptr,
                       (rng() % 2) ? mask.data() : nullptr};
        hbo_lut lut;
        CHECK(hbo_lut_build(&Q, &lut) == 0);
        const uint32_t ns = 1 + (uint32_t)(rng() % 60);
        std::vector<uint64_t> offs(ns + 1, 0);
        std::vector<uint8_t> res;
        for (uint32_t k = 0; k < ns; ++k) {
            const size_t L = rng() % 400;
            for (size_t i = 0; i < L; ++i) {
                // homologous stretches from the query make HSPs
                const bool copy = (i % 97) < 40 && (rng() % 10) < 8;
                res.push_back(copy ? q[(i * 7 + k) % qlen] : (uint8_t)(1 + rng() % 22));
            }
            offs[k + 1] = res.size();
        }
        res.resize(res.size() + 64, 0);
        hbo_hsp *hsps = nullptr;
        std::vector<uint32_t> cnt(3 * ns);
        const int64_t nh =
            hbo_scan_db(&Q, &lut, res.data(), offs.data(), ns, threads, &hsps, cnt.data());
        CHECK(nh >= 0);
        if (nh >= 0) {
            hbo_gap_params gp = {11, 1, (int32_t)(rng() % 80)};
            hbo_ghsp *g = nullptr;
            std::vector<int32_t> gc(2 * ns);
            int32_t gcut[2] = {25, 30};
            const int64_t ng =
                hbo_gapped_db(&Q, res.data(), offs.data(), ns, hsps, (uint64_t)nh, &gp, gcut,
                              threads, &g, gc.data());
            CHECK(ng >= 0 && ng <= nh);
            hbo_free(g);
        }
        hbo_free(hsps);
        hbo_lut_free(&lut);
        // the DP on random operands, both directions, clamped and huge X
        const int32_t M = (int32_t)(rng() % 40), N = (int32_t)(rng() % 40);
        std::vector<uint8_t> A(M + 2), B(N + 2);
        for (auto &x : A) x = (uint8_t)(1 + rng() % 22);
        for (auto &x : B) x = (uint8_t)(1 + rng() % 22);
        int32_t ao, bo;
        for (int32_t X : {0, 12, 38, 100000}) {
            const int32_t s0 = hbo_semigapped(hbo_blosum62(), A.data(), B.data(), M, N, 0, 11, 1,
                                              X, &ao, &bo);
            CHECK(s0 >= 0 && ao <= M && bo <= N);
            const int32_t s1 = hbo_semigapped(hbo_blosum62(), A.data(), B.data(), M, N, 1, 11, 1,
                                              X, &ao, &bo);
            CHECK(s1 >= 0 && ao <= M && bo <= N);
        }
    }
}

// .hDb writer against the oracle's writer and reader; .hDbI text round trip,
// then random corruptions of it through the parser.
static void hdb(std::mt19937_64 &rng, int rounds)
{
    for (int r = 0; r < rounds; ++r) {
        const int32_t n = (int32_t)(rng() % 300);
        std::vector<uint64_t> lens(n);
        for (auto &l : lens) l = rng() % 200;
        std::sort(lens.begin(), lens.end());
        std::vector<uint64_t> offs(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) offs[i + 1] = offs[i] + lens[i];
        std::vector<uint8_t> res(offs[n] + 1);
        for (auto &c : res) c = (uint8_t)(1 + rng() % 24);
        const int32_t bx = 1 + (int32_t)(rng() % 8), tx = 1 + (int32_t)(rng() % 70);
        int32_t ng = 0;
        uint64_t bytes = 0;
        CHECK(hb_hdb_layout(offs.data(), n, bx, tx, nullptr, 0, &ng, &bytes) == HB_OK);
        std::vector<int32_t> gl(ng + 1), gl_o(ng + 1);
        CHECK(hb_hdb_layout(offs.data(), n, bx, tx, gl.data(), ng, &ng, &bytes) == HB_OK);
        uint64_t bytes_o = 0;
        CHECK(hbo_hdb_layout(offs.data(), n, bx, tx, gl_o.data(), &bytes_o) == 0);
        CHECK(bytes == bytes_o);
        std::vector<uint8_t> a(bytes + 4), b(bytes + 4);
        CHECK(hb_hdb_pack(res.data(), offs.data(), n, bx, tx, gl.data(), ng, a.data(), bytes) ==
              HB_OK);
        CHECK(hbo_hdb_pack(res.data(), offs.data(), n, bx, tx, gl_o.data(), b.data(), bytes) == 0);
        CHECK(std::memcmp(a.data(), b.data(), bytes) == 0);
        std::vector<uint64_t> o2(n + 1);
        std::vector<uint8_t> r2(offs[n] + 1);
        CHECK(hbo_hdb_unpack(a.data(), bytes, n, bx, tx, gl.data(), r2.data(), o2.data()) == 0);
        CHECK(o2 == offs);
        // index text with two volumes
        hb_hdb_volume v[2] = {};
        for (int k = 0; k < 2; ++k) {
            v[k].name = k ? "db.01.hDb" : "db.00.hDb";
            v[k].bytes = bytes;
            v[k].volume_seqs = n + k;
            v[k].gpu_seqs = n;
            v[k].n_groups = ng;
            v[k].group_len = gl.data();
        }
        hb_hdb_index ix = {bx, tx, 99, 2};
        uint64_t len = 0;
        CHECK(hb_hdbi_format(&ix, v, nullptr, 0, &len) == HB_OK);
        std::string t(len, '\0');
        CHECK(hb_hdbi_format(&ix, v, &t[0], len, &len) == HB_OK);
        for (int m = 0; m < 20; ++m) {
            std::string u = t;
            const int edits = m == 0 ? 0 : 1 + (int)(rng() % 4);
            for (int e = 0; e < edits && !u.empty(); ++e) {
                const size_t at = rng() % u.size();
                switch (rng() % 4) {
                case 0: u.erase(at, 1 + rng() % 8); break;
                case 1: u.insert(at, 1, "0123456789#\n-x"[rng() % 15]); break;
                case 2: u[at] = "0123456789#\n-x "[rng() % 16]; break;
                default: u.resize(at); break;
                }
            }
            hb_hdb_index jx;
            uint64_t ngl = 0, eo = 0;
            const int rc = hb_hdbi_parse(u.data(), u.size(), &jx, nullptr, 0, nullptr, 0, &ngl, &eo);
            if (m == 0) CHECK(rc == HB_OK && jx.n_volumes == 2 && ngl == 2 * (uint64_t)ng);
            if (rc == HB_ERR_FORMAT) CHECK(eo <= u.size());
            if (rc != HB_OK || jx.n_volumes > 64 || ngl > (1u << 20)) continue;
            std::vector<hb_hdb_volume> vs(jx.n_volumes + 1);
            std::vector<int32_t> g(ngl + 1);
            CHECK(hb_hdbi_parse(u.data(), u.size(), &jx, vs.data(), jx.n_volumes, g.data(), ngl,
                                &ngl, &eo) == HB_OK);
            if (jx.n_volumes > 0 && ngl > 0)   // too little room for the group lengths
                CHECK(hb_hdbi_parse(u.data(), u.size(), &jx, vs.data(), jx.n_volumes, g.data(),
                                    ngl - 1, &ngl, &eo) == HB_ERR_CAPACITY);
            if (m == 0)
                for (int k = 0; k < 2; ++k)
                    CHECK(vs[k].n_groups == ng && vs[k].name_len == 9 &&
                          std::memcmp(vs[k].group_len, gl.data(), 4 * (size_t)ng) == 0);
        }
    }
}

// Options texts with random lines and truncations, splits of random volumes
// and random balancer inputs, product against the oracle.
static void balance(std::mt19937_64 &rng, int rounds)
{
    static const char *tok[] = {"0", "1", "-1", "0.5", "12", "x", "", "1e9", "99999999999", "nan"};
    for (int r = 0; r < rounds; ++r) {
        std::string t;
        for (int i = 0; i < 9; ++i) {
            t += "#c\n";
            t += tok[rng() % 10];
            t += "\n";
        }
        if (rng() % 4 == 0) t.resize(rng() % (t.size() + 1));
        hb_runtime_options a;
        hbo_options b;
        const int ra = hb_runtime_options_parse(t.data(), t.size(), &a);
        const int rb = hbo_options_parse(t.data(), t.size(), &b);
        CHECK((ra == HB_OK) == (rb == 0));
        // a volume's groups and a ratio
        const int32_t bx = 1 + (int32_t)(rng() % 5), tx = 1 + (int32_t)(rng() % 70);
        const int32_t G = bx * tx, ng = 1 + (int32_t)(rng() % 10);
        const int32_t n = (ng - 1) * G + 1 + (int32_t)(rng() % G);
        std::vector<int32_t> gl(ng);
        for (auto &x : gl) x = 16 * (int32_t)(rng() % 40);
        std::sort(gl.begin(), gl.end());
        float r1 = (rng() % 5 == 0) ? -1.0f : (float)(rng() % 1200) / 1000.0f, r2 = r1;
        int32_t f1 = (int32_t)(rng() % 2), f2 = f1;
        hb_hdb_share sh;
        int32_t s2, g2;
        uint64_t b2;
        CHECK(hb_hdb_split(gl.data(), ng, n, 12345, bx, tx, 4, &r1, &f1, &sh) == HB_OK);
        CHECK(hbo_split(gl.data(), ng, n, 12345, tx, G, 4, &r2, &f2, &s2, &g2, &b2) == 0);
        CHECK(sh.gpu_seqs == s2 && sh.n_groups == g2 && sh.bytes == b2 && f1 == f2);
        CHECK(std::memcmp(&r1, &r2, sizeof r1) == 0);
        hb_runtime_options o;
        hb_runtime_options_default(&o);
        hbo_options oo;
        hbo_options_default(&oo);
        hb_balance_in in = {};
        in.volume = 1 + (uint32_t)(rng() % 9);
        in.ratio = (float)(rng() % 1000) / 1000.0f;
        in.cpu_threads = 1 + (uint32_t)(rng() % 64);
        in.gpu_cards = 1;
        for (auto &x : in.volume_bytes) x = 1 + rng() % 1000000;
        for (auto &x : in.ratio_ring) x = (float)(1 + rng() % 1000) / 1000.0f;
        for (auto &x : in.hsp_ring) x = (int64_t)(rng() % 100000);
        in.gapped = (int32_t)(rng() % 2);
        in.query_length = (int32_t)(rng() % 10000);
        in.avg_query_length = (uint32_t)(rng() % 3000);
        in.gpu_seconds = (double)(rng() % 1000) / 100.0;
        in.cpu_seconds = (double)(rng() % 1000) / 100.0;
        float x1 = 0, x2 = 0;
        CHECK(hb_balance_next_ratio(&o, &in, &x1) == HB_OK);
        CHECK(hbo_next_ratio(&oo, in.volume, in.ratio, in.cpu_threads, in.gpu_cards,
                             in.volume_bytes, in.ratio_ring, in.hsp_ring, in.gapped,
                             in.query_length, in.avg_query_length, in.gpu_seconds,
                             in.cpu_seconds, &x2) == 0);
        CHECK(x1 == x2);
    }
}

// SEG masks of random queries with low-complexity runs and other letters.
static void seg(std::mt19937_64 &rng, int rounds)
{
    for (int r = 0; r < rounds; ++r) {
        std::vector<uint8_t> q;
        const int parts = 1 + (int)(rng() % 5);
        for (int k = 0; k < parts; ++k) {
            const int n = (int)(rng() % 120);
            for (int i = 0; i < n; ++i) q.push_back((uint8_t)(1 + rng() % 27));
            const int m = (int)(rng() % 50);
            const uint8_t a = (uint8_t)(1 + rng() % 20), b = (uint8_t)(1 + rng() % 20);
            for (int i = 0; i < m; ++i) q.push_back(i % 2 ? a : b);
        }
        const int32_t len = (int32_t)q.size();
        const int32_t half = len / 2;
        const int32_t off[2] = {0, half}, cl[2] = {half, len - half};
        hb_seg_params p;
        hb_seg_params_default(&p);
        p.window = 1 + (int32_t)(rng() % 20);
        p.maxtrim = (int32_t)(rng() % 120);
        std::vector<uint8_t> m1(len + 1), m2(len + 1);
        CHECK(hb_seg_mask(q.data(), len, 2, off, cl, &p, m1.data()) == HB_OK);
        CHECK(hbo_seg_mask(q.data(), len, 2, off, cl, p.window, p.locut, p.hicut, p.maxbogus,
                           p.maxtrim, m2.data()) == 0);
        CHECK(std::memcmp(m1.data(), m2.data(), (size_t)len) == 0);
    }
}

// Volume cut and length order of random sequence sets (permutation, each
// volume shortest first, cut sizes as the rule gives them), written out
// with hb_hdb_write and read back through the index.
static void volumes(std::mt19937_64 &rng, int rounds, const std::string &dir)
{
    for (int it = 0; it < rounds; ++it) {
        const int32_t n = (int32_t)(rng() % 400);
        std::vector<uint64_t> offs(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) offs[i + 1] = offs[i] + rng() % 300;
        const uint64_t maxb = 2 + rng() % 20000, maxl = rng() % 2 ? 0 : 1 + rng() % 10000;
        std::vector<int32_t> order(std::max(n, 1)), vs(n + 1);
        int32_t nv = 0;
        CHECK(hb_db_sort_volumes(offs.data(), n, maxb, maxl, order.data(), vs.data(), n + 1,
                                 &nv) == HB_OK);
        std::vector<char> seen(n, 0);
        int32_t first = 0;
        for (int32_t v = 0; v < nv; ++v) {
            CHECK(vs[v] >= 1);
            for (int32_t k = first; k < first + vs[v]; ++k) {
                CHECK(order[k] >= first && order[k] < first + vs[v] && !seen[order[k]]);
                seen[order[k]] = 1;
                if (k > first)
                    CHECK(offs[order[k] + 1] - offs[order[k]] >=
                          offs[order[k - 1] + 1] - offs[order[k - 1]]);
            }
            first += vs[v];
        }
        CHECK(first == n);
        if (n == 0 || it % 10) continue;
        std::vector<uint8_t> res(offs[n]);
        for (auto &r : res) r = (uint8_t)(1 + rng() % 24);
        const std::string pre = dir + "/san_vol";
        CHECK(hb_hdb_write(pre.c_str(), res.data(), offs.data(), n, order.data(), vs.data(), nv,
                           1 + (int32_t)(rng() % 3), 1 + (int32_t)(rng() % 64)) == HB_OK);
        FILE *f = std::fopen((pre + ".hDbI").c_str(), "rb");
        CHECK(f != nullptr);
        if (!f) continue;
        std::string t;
        char b[4096];
        size_t got;
        while ((got = std::fread(b, 1, sizeof(b), f)) > 0) t.append(b, got);
        std::fclose(f);
        hb_hdb_index ix;
        uint64_t ngl = 0, eo = 0;
        CHECK(hb_hdbi_parse(t.data(), t.size(), &ix, nullptr, 0, nullptr, 0, &ngl, &eo) == HB_OK);
        CHECK(ix.n_volumes == nv);
        // a bad order is refused
        if (n > 1) {
            std::vector<int32_t> bad(order.begin(), order.begin() + n);
            bad[1] = bad[0];
            CHECK(hb_hdb_write(pre.c_str(), res.data(), offs.data(), n, bad.data(), vs.data(), nv,
                               1, 32) == HB_ERR_ARG);
        }
    }
}

// The C++ co-scheduled loop over a written database with fake device steps:
// random fits, out-of-memory launches, device and CPU times, a failing CPU
// search; the order invariants hold and every started share is fetched or
// abandoned.
struct FakeDev {
    std::mt19937_64 *rng;
    int launched = 0, fetched = 0, abandoned = 0, fail_cpu_at = -1;
};
static int fake_fits(void *u, int32_t, const hb_hdb_share *sh, int32_t, uint64_t *need,
                     int32_t *fits)
{
    FakeDev *d = static_cast<FakeDev *>(u);
    *need = sh->bytes * 3 + 64;
    *fits = (*d->rng)() % 5 != 0;
    return HB_OK;
}
static int fake_launch(void *u, int32_t, const uint8_t *hdb, const hb_hdb_share *sh, int32_t,
                       int32_t, const int32_t *)
{
    FakeDev *d = static_cast<FakeDev *>(u);
    volatile uint8_t x = sh->bytes ? hdb[sh->bytes - 1] : 0;   // the share was read whole
    (void)x;
    if ((*d->rng)() % 7 == 0) return HB_ERR_NOMEM;
    ++d->launched;
    return HB_OK;
}
static int fake_fetch(void *u, int32_t, int64_t, double *secs, uint64_t *n_hsp)
{
    FakeDev *d = static_cast<FakeDev *>(u);
    ++d->fetched;
    *secs = 0.001 * (double)((*d->rng)() % 100);
    *n_hsp = (*d->rng)() % 1000;
    return HB_OK;
}
static void fake_abandon(void *u) { ++static_cast<FakeDev *>(u)->abandoned; }
static int fake_cpu(void *u, int32_t v, int64_t lo, int64_t hi)
{
    FakeDev *d = static_cast<FakeDev *>(u);
    CHECK(lo <= hi);
    return v == d->fail_cpu_at ? 1 : 0;
}

static void cosched(std::mt19937_64 &rng, int rounds, const std::string &dir)
{
    for (int it = 0; it < rounds; ++it) {
        const int32_t n = 200 + (int32_t)(rng() % 3000);
        std::vector<uint64_t> offs(n + 1, 0);
        for (int32_t i = 0; i < n; ++i) offs[i + 1] = offs[i] + 1 + rng() % 200;
        std::vector<uint8_t> res(offs[n]);
        for (auto &r : res) r = (uint8_t)(1 + rng() % 24);
        std::vector<int32_t> order(n), vs(n + 1);
        int32_t nv = 0;
        CHECK(hb_db_sort_volumes(offs.data(), n, 20000 + rng() % 100000, 0, order.data(), vs.data(),
                                 n + 1, &nv) == HB_OK);
        const std::string pre = dir + "/san_cos";
        CHECK(hb_hdb_write(pre.c_str(), res.data(), offs.data(), n, order.data(), vs.data(), nv, 2,
                           16) == HB_OK);
        FakeDev d;
        d.rng = &rng;
        d.fail_cpu_at = rng() % 3 == 0 ? (int)(rng() % nv) : -1;
        hb_cosched_steps st = {&d, fake_fits, fake_launch, fake_fetch, fake_abandon};
        hb_runtime_options opts;
        hb_runtime_options_default(&opts);
        std::vector<hb_volume_report> rep(nv);
        int32_t got = 0;
        const int rc = hb_cosched_run((pre + ".hDbI").c_str(), &opts, 1 + (int32_t)(rng() % 16), 0,
                                      300, 300, &st, fake_cpu, &d, rep.data(), nv, &got);
        CHECK(got == nv);
        CHECK(d.fetched + d.abandoned == d.launched);
        // after a volume whose ratio is 0 every later one goes to the CPU
        bool zero = false;
        for (int32_t v = 0; v < nv; ++v) {
            if (zero) CHECK(rep[v].gpu_seqs == 0 && rep[v].ratio == 0.f);
            zero = zero || rep[v].ratio == 0.f;
        }
        if (d.fail_cpu_at < 0) {
            CHECK(rc == HB_OK);
            for (int32_t v = 0; v + 1 < nv; ++v)
                if (rep[v + 1].launch_seq) CHECK(rep[v + 1].launch_seq < rep[v].cpu_seq);
        } else {
            CHECK(rc == HB_ERR_STATE);
        }
    }
}

int main(int argc, char **argv)
{
    const bool threads_only = argc > 1 && std::strcmp(argv[1], "threads") == 0;
    const std::string dir = argc > 1 && !threads_only ? argv[1] : ".";
    std::mt19937_64 rng(20261202);
    if (threads_only) {
        oracle(rng, 30, 4);
    } else {
        fasta(rng, 3000);
        translate(rng, 300);
        oracle(rng, 60, 1);
        hdb(rng, 300);
        balance(rng, 2000);
        seg(rng, 300);
        volumes(rng, 300, dir);
        cosched(rng, 40, dir);
    }
    std::fprintf(stderr, "%s: %d failed checks\n", threads_only ? "threads" : "all", fails);
    return fails ? 1 : 0;
}
