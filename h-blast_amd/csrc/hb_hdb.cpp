// H-BLAST GPU database volumes (.hDb) and their text index (.hDbI): layout,
// writer and index parser/formatter (host code, no HIP).  The device
// unpacking of a volume is hb_db_upload_hdb (hb_seed.cpp, kernels in
// hb_hdb.hip).
//
// Format, as CreateH_BLAST_database writes it
// (ncbi_blast_files/functions.cpp:785-1165):
//   - the volume is sorted by ascending length; its first
//     floor(floor(99.9% n) / 32) * 32 sequences go to the file (:842-843);
//   - groups of blocksx * threadsx consecutive sequences; a group's length
//     is its last (longest) member's length rounded up to 16 (:844-903,
//     Round2Multiple, H-BLAST/h_blast.c:11-17); a full group occupies
//     group size * length bytes, the last one whole blocks of threadsx
//     sequences (:906-911);
//   - in a block, the t-th sequence's residues 4i..4i+3 sit at byte
//     (t + i * threadsx) * 4 of the block (:1058-1066); the rest is zero.
//   - the index text: a comment line before every value (:257-309).
// An empty sequence writes nothing (the reference stores a byte of it
// that does not exist, :1058-1066).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hb_seed.h"

namespace {

constexpr int32_t kGridMax = 65536;   // -gpu_threads / -gpu_blocks limit (blast_args.cpp:2726,2733)
constexpr uint64_t kSeqMax = 1ull << 28;

int32_t round_up16(uint64_t len) { return (int32_t)((len + 15) & ~15ull); }

bool grid_ok(int32_t blocksx, int32_t threadsx)
{
    return blocksx >= 1 && blocksx <= kGridMax && threadsx >= 1 && threadsx <= kGridMax;
}

}  // namespace

extern "C" int32_t hb_hdb_gpu_seqs(int32_t volume_seqs)
{
    if (volume_seqs <= 0) return 0;
    // the reference computes (99.9 * n) / 100 in double and truncates
    const int32_t share = (int32_t)((99.9 * (double)volume_seqs) / 100.0);
    return share - share % 32;
}

extern "C" int hb_hdb_layout(const uint64_t *offsets, int32_t n_seqs, int32_t blocksx,
                             int32_t threadsx, int32_t *group_len, int32_t group_cap,
                             int32_t *n_groups, uint64_t *bytes)
{
    if (!offsets || !n_groups || !bytes || n_seqs < 0 || !grid_ok(blocksx, threadsx))
        return HB_ERR_ARG;
    uint64_t prev = 0;
    for (int32_t i = 0; i < n_seqs; ++i) {
        if (offsets[i + 1] < offsets[i]) return HB_ERR_ARG;
        const uint64_t len = offsets[i + 1] - offsets[i];
        if (len > kSeqMax || len < prev) return HB_ERR_ARG;   // too long / not sorted
        prev = len;
    }
    const uint64_t G = (uint64_t)blocksx * (uint64_t)threadsx;
    const uint64_t ng = ((uint64_t)n_seqs + G - 1) / G;
    *n_groups = (int32_t)ng;
    uint64_t total = 0;
    for (uint64_t g = 0; g < ng; ++g) {
        const uint64_t first = g * G;
        const uint64_t members = std::min<uint64_t>(G, (uint64_t)n_seqs - first);
        const uint64_t last = first + members - 1;
        const int32_t L = round_up16(offsets[last + 1] - offsets[last]);
        const uint64_t slots = (members + threadsx - 1) / threadsx * (uint64_t)threadsx;
        total += slots * (uint64_t)L;
        if (group_len) {
            if ((int64_t)g >= group_cap) return HB_ERR_CAPACITY;
            group_len[g] = L;
        }
    }
    *bytes = total;
    return HB_OK;
}

extern "C" int hb_hdb_pack(const uint8_t *residues, const uint64_t *offsets, int32_t n_seqs,
                           int32_t blocksx, int32_t threadsx, const int32_t *group_len,
                           int32_t n_groups, uint8_t *out, uint64_t out_bytes)
{
    if (!offsets || !group_len || (n_seqs > 0 && (!residues || !out)) || n_seqs < 0 ||
        !grid_ok(blocksx, threadsx))
        return HB_ERR_ARG;
    int32_t ng = 0;
    uint64_t need = 0;
    std::vector<int32_t> lens;
    {
        const uint64_t G = (uint64_t)blocksx * (uint64_t)threadsx;
        lens.resize(((uint64_t)n_seqs + G - 1) / G);
        const int rc = hb_hdb_layout(offsets, n_seqs, blocksx, threadsx, lens.data(),
                                     (int32_t)lens.size(), &ng, &need);
        if (rc != HB_OK) return rc;
    }
    if (ng != n_groups || need != out_bytes) return HB_ERR_ARG;
    for (int32_t g = 0; g < ng; ++g)
        if (lens[g] != group_len[g]) return HB_ERR_ARG;
    if (out_bytes) std::memset(out, 0, out_bytes);
    const uint64_t T = (uint64_t)threadsx, G = (uint64_t)blocksx * T;
    uint64_t group_base = 0;
    for (int32_t g = 0; g < ng; ++g) {
        const uint64_t L = (uint64_t)group_len[g];
        const uint64_t first = (uint64_t)g * G;
        const uint64_t members = std::min<uint64_t>(G, (uint64_t)n_seqs - first);
        for (uint64_t m = 0; m < members; ++m) {
            const uint64_t c = first + m;
            uint8_t *blk = out + group_base + (m / T) * T * L;
            const uint64_t t = m % T;
            const uint8_t *s = residues + offsets[c];
            const uint64_t len = offsets[c + 1] - offsets[c];
            for (uint64_t a = 0; a < len; a += 4)
                std::memcpy(blk + (t + (a / 4) * T) * 4, s + a, std::min<uint64_t>(4, len - a));
        }
        group_base += members == G ? G * L : (members + T - 1) / T * T * L;
    }
    return HB_OK;
}

extern "C" int hb_hdbi_format(const hb_hdb_index *index, const hb_hdb_volume *vols, char *text,
                              uint64_t cap, uint64_t *len)
{
    if (!index || !len || index->n_volumes < 0 || (index->n_volumes > 0 && !vols))
        return HB_ERR_ARG;
    std::string s;
    auto put = [&](const char *comment, const std::string &value) {
        s += comment;
        s += '\n';
        s += value;
        s += '\n';
    };
    put("#Number of blocksx", std::to_string(index->blocksx));
    put("#Number of threadsx", std::to_string(index->threadsx));
    put("#Percentage", std::to_string(index->percentage));
    put("#Number of volumes", std::to_string(index->n_volumes));
    for (int32_t v = 0; v < index->n_volumes; ++v) {
        const hb_hdb_volume &V = vols[v];
        if (!V.name || V.n_groups < 0 || (V.n_groups > 0 && !V.group_len)) return HB_ERR_ARG;
        const std::string name = V.name_len ? std::string(V.name, V.name_len) : std::string(V.name);
        if (name.find('\n') != std::string::npos) return HB_ERR_ARG;
        put("#Volume name", name);
        put("#Volume size in bytes", std::to_string(V.bytes));
        put("#Volume sequences", std::to_string(V.volume_seqs));
        put("#Number of sequences in GPU volume", std::to_string(V.gpu_seqs));
        put("#Store limit", std::to_string(V.store_limit));
        put("#Break limit", std::to_string(V.break_limit));
        put("#Group number", std::to_string(V.n_groups));
        s += "#Length of each group\n";
        for (int32_t g = 0; g < V.n_groups; ++g) {
            s += std::to_string(V.group_len[g]);
            s += '\n';
        }
        put("#Length of the rest sequences in CPU database", std::to_string(V.cpu_rest_bytes));
    }
    *len = s.size();
    if (!text) return HB_OK;
    if (cap < s.size()) return HB_ERR_CAPACITY;
    std::memcpy(text, s.data(), s.size());
    return HB_OK;
}

namespace {

// Line reader over the index text; a line ends at '\n' (a '\r' before it
// is dropped) or at the end of the text.
struct Lines {
    const char *t;
    uint64_t n, pos = 0, line_start = 0;
    bool next(const char *&b, uint64_t &len)
    {
        if (pos >= n) return false;
        line_start = pos;
        const char *nl = (const char *)std::memchr(t + pos, '\n', n - pos);
        const uint64_t end = nl ? (uint64_t)(nl - t) : n;
        b = t + pos;
        len = end - pos;
        if (len && b[len - 1] == '\r') --len;
        pos = nl ? end + 1 : n;
        return true;
    }
};

bool parse_i64(const char *b, uint64_t len, int64_t lo, int64_t hi, int64_t &v)
{
    while (len && (b[len - 1] == ' ' || b[len - 1] == '\t')) --len;
    uint64_t i = 0;
    while (i < len && (b[i] == ' ' || b[i] == '\t')) ++i;
    bool neg = false;
    if (i < len && (b[i] == '-' || b[i] == '+')) neg = b[i++] == '-';
    if (i == len) return false;
    unsigned long long x = 0;
    for (; i < len; ++i) {
        if (b[i] < '0' || b[i] > '9') return false;
        if (x > (~0ull - 9) / 10) return false;
        x = x * 10 + (unsigned)(b[i] - '0');
    }
    if (x > (1ull << 62)) return false;
    v = neg ? -(int64_t)x : (int64_t)x;
    return v >= lo && v <= hi;
}

}  // namespace

extern "C" int hb_hdbi_parse(const char *text, uint64_t len, hb_hdb_index *index,
                             hb_hdb_volume *vols, int32_t vol_cap, int32_t *group_len,
                             uint64_t group_cap, uint64_t *n_group_len, uint64_t *err_off)
{
    if (!text || !index || !n_group_len) return HB_ERR_ARG;
    Lines L{text, len};
    const char *b = nullptr;
    uint64_t n = 0;
    auto fail = [&]() {
        if (err_off) *err_off = L.line_start;
        return HB_ERR_FORMAT;
    };
    // a comment line, then the value line
    auto value = [&](int64_t lo, int64_t hi, int64_t &v) {
        if (!L.next(b, n) || n == 0 || b[0] != '#') return false;
        return L.next(b, n) && parse_i64(b, n, lo, hi, v);
    };
    const int64_t I32 = INT32_MAX;
    int64_t bx, tx, pc, nv;
    if (!value(1, kGridMax, bx) || !value(1, kGridMax, tx) || !value(-I32, I32, pc) ||
        !value(0, I32, nv))
        return fail();
    index->blocksx = (int32_t)bx;
    index->threadsx = (int32_t)tx;
    index->percentage = (int32_t)pc;
    index->n_volumes = (int32_t)nv;
    const bool fill = vols && group_len;
    if (fill && vol_cap < nv) return HB_ERR_CAPACITY;
    uint64_t ng_total = 0;
    for (int64_t v = 0; v < nv; ++v) {
        hb_hdb_volume V{};
        if (!L.next(b, n) || n == 0 || b[0] != '#' || !L.next(b, n) || n == 0) return fail();
        V.name = b;
        V.name_len = (uint32_t)n;
        int64_t bytes, vs, gs, sl, bl, ng, rest;
        if (!value(0, INT64_MAX, bytes) || !value(0, I32, vs) || !value(0, I32, gs) ||
            !value(0, I32, sl) || !value(0, I32, bl) || !value(0, I32, ng))
            return fail();
        if (!L.next(b, n) || n == 0 || b[0] != '#') return fail();
        if (fill && ng_total + (uint64_t)ng > group_cap) return HB_ERR_CAPACITY;
        for (int64_t g = 0; g < ng; ++g) {
            int64_t x;
            if (!L.next(b, n) || !parse_i64(b, n, 0, I32, x)) return fail();
            if (fill) group_len[ng_total + g] = (int32_t)x;
        }
        if (!value(0, INT64_MAX, rest)) return fail();
        V.bytes = (uint64_t)bytes;
        V.volume_seqs = (int32_t)vs;
        V.gpu_seqs = (int32_t)gs;
        V.store_limit = (int32_t)sl;
        V.break_limit = (int32_t)bl;
        V.n_groups = (int32_t)ng;
        V.group_len = fill ? group_len + ng_total : nullptr;
        V.cpu_rest_bytes = (uint64_t)rest;
        if (fill) vols[v] = V;
        ng_total += (uint64_t)ng;
    }
    // only blank lines may follow
    while (L.next(b, n))
        if (n) return fail();
    *n_group_len = ng_total;
    return HB_OK;
}

// ------------------------------------------------------------------ sorted volumes
// H-BLAST's makeblastdb with -sort_volumes (CWriteDB_Volume::WriteSequenceSort,
// ncbi_blast_files/writedb_volume.cpp:255-513) fills a volume with the input
// sequences in input order until the next one would overfill its files, then
// sorts the volume with std::sort by descending length (compare, :159-164) and
// writes it from the back, so shortest first.  The sequence file's rule
// (CWriteDB_SequenceFile::CanFit and UpdateSize,
// ncbi_blast_files/writedb_files.hpp:435-494): the file starts at offset 1,
// each sequence adds its length + 1 (a NUL after it), a volume's first
// sequence always fits, and a later one fits while offset + length <
// max_file_sz and, with a letter limit, letters + length <= the limit.
extern "C" int hb_db_sort_volumes(const uint64_t *offsets, int32_t n_seqs,
                                  uint64_t max_file_bytes, uint64_t max_letters, int32_t *order,
                                  int32_t *volume_seqs, int32_t volume_cap, int32_t *n_volumes)
{
    if (!offsets || n_seqs < 0 || !n_volumes || max_file_bytes < 2 || (n_seqs > 0 && !order) ||
        volume_cap < 0)
        return HB_ERR_ARG;
    std::vector<int32_t> cut;
    uint64_t off = 1, letters = 0;
    int32_t cnt = 0;
    for (int32_t i = 0; i < n_seqs; ++i) {
        if (offsets[i + 1] < offsets[i]) return HB_ERR_ARG;
        const uint64_t len = offsets[i + 1] - offsets[i];
        if (cnt > 0 && (off + len >= max_file_bytes || (max_letters && letters + len > max_letters))) {
            cut.push_back(cnt);
            off = 1;
            letters = 0;
            cnt = 0;
        }
        off += len + 1;
        letters += len;
        ++cnt;
    }
    if (cnt > 0) cut.push_back(cnt);
    *n_volumes = (int32_t)cut.size();
    if (volume_seqs) {
        if (volume_cap < (int32_t)cut.size()) return HB_ERR_CAPACITY;
        std::copy(cut.begin(), cut.end(), volume_seqs);
    }
    // the reference's sort on the lengths alone: std::sort moves elements by
    // the comparisons' outcomes only, so the permutation is the one it
    // applies to its sequence objects
    struct Item {
        uint64_t len;
        int32_t idx;
    };
    std::vector<Item> v;
    int32_t first = 0;
    for (int32_t c : cut) {
        v.clear();
        for (int32_t i = first; i < first + c; ++i) v.push_back(Item{offsets[i + 1] - offsets[i], i});
        std::sort(v.begin(), v.end(), [](const Item &a, const Item &b) { return a.len > b.len; });
        for (int32_t k = 0; k < c; ++k) order[first + k] = v[(size_t)(c - 1 - k)].idx;
        first += c;
    }
    return HB_OK;
}

// Writes an H-BLAST GPU database from length-sorted volumes: the database's
// k-th sequence is input sequence order[k] (order NULL: k), volume_seqs
// counts the volumes' sequences; <prefix>.hDb for one volume, else
// <prefix>.NN.hDb per volume (CreateH_BLAST_database's names,
// functions.cpp:1089-1127), and the index <prefix>.hDbI.
extern "C" int hb_hdb_write(const char *prefix, const uint8_t *residues, const uint64_t *offsets,
                            int32_t n_seqs, const int32_t *order, const int32_t *volume_seqs,
                            int32_t n_volumes, int32_t blocksx, int32_t threadsx)
{
    if (!prefix || !offsets || !volume_seqs || n_volumes < 1 || n_seqs < 0 ||
        (n_seqs > 0 && !residues) || !grid_ok(blocksx, threadsx))
        return HB_ERR_ARG;
    if (order) {   // a permutation of 0 .. n_seqs - 1
        std::vector<uint8_t> seen((size_t)n_seqs, 0);
        for (int32_t k = 0; k < n_seqs; ++k) {
            if (order[k] < 0 || order[k] >= n_seqs || seen[(size_t)order[k]]) return HB_ERR_ARG;
            seen[(size_t)order[k]] = 1;
        }
    }
    for (int32_t i = 0; i < n_seqs; ++i)
        if (offsets[i + 1] < offsets[i]) return HB_ERR_ARG;
    int64_t total = 0;
    for (int32_t k = 0; k < n_volumes; ++k) {
        if (volume_seqs[k] < 0) return HB_ERR_ARG;
        total += volume_seqs[k];
    }
    if (total != n_seqs) return HB_ERR_ARG;
    const std::string pre(prefix);
    const size_t slash = pre.find_last_of('/');
    const std::string stem = slash == std::string::npos ? pre : pre.substr(slash + 1);
    const std::string dir = slash == std::string::npos ? std::string() : pre.substr(0, slash + 1);
    std::vector<std::string> names((size_t)n_volumes);
    std::vector<std::vector<int32_t>> groups((size_t)n_volumes);
    std::vector<hb_hdb_volume> vols((size_t)n_volumes);
    std::vector<uint8_t> buf, gathered;
    std::vector<uint64_t> o;
    int32_t first = 0;
    for (int32_t k = 0; k < n_volumes; ++k) {
        const int32_t vs = volume_seqs[k];
        o.assign((size_t)vs + 1, 0);
        const uint8_t *r = nullptr;
        if (order) {   // the volume's sequences gathered in database order
            for (int32_t i = 0; i < vs; ++i) {
                const int32_t src = order[first + i];
                o[(size_t)i + 1] = o[(size_t)i] + (offsets[src + 1] - offsets[src]);
            }
            gathered.resize(o[(size_t)vs]);
            for (int32_t i = 0; i < vs; ++i) {
                const int32_t src = order[first + i];
                if (o[(size_t)i + 1] > o[(size_t)i])
                    std::memcpy(gathered.data() + o[(size_t)i], residues + offsets[src],
                                o[(size_t)i + 1] - o[(size_t)i]);
            }
            r = gathered.data();
        } else {
            for (int32_t i = 0; i <= vs; ++i) o[(size_t)i] = offsets[first + i] - offsets[first];
            r = residues ? residues + offsets[first] : nullptr;
        }
        const int32_t gs = hb_hdb_gpu_seqs(vs);
        int32_t ng = 0;
        uint64_t bytes = 0;
        int rc = hb_hdb_layout(o.data(), gs, blocksx, threadsx, nullptr, 0, &ng, &bytes);
        if (rc) return rc;
        groups[(size_t)k].resize((size_t)std::max(ng, 1));
        rc = hb_hdb_layout(o.data(), gs, blocksx, threadsx, groups[(size_t)k].data(), ng, &ng,
                           &bytes);
        if (rc) return rc;
        buf.assign(bytes, 0);
        if (bytes) {
            rc = hb_hdb_pack(r, o.data(), gs, blocksx, threadsx, groups[(size_t)k].data(), ng,
                             buf.data(), bytes);
            if (rc) return rc;
        }
        char digits[16];
        std::snprintf(digits, sizeof(digits), ".%02d", k);
        names[(size_t)k] = stem + (n_volumes > 1 ? std::string(digits) : std::string()) + ".hDb";
        FILE *f = std::fopen((dir + names[(size_t)k]).c_str(), "wb");
        if (!f) return HB_ERR_ARG;
        const size_t put = bytes ? std::fwrite(buf.data(), 1, bytes, f) : 0;
        if (std::fclose(f) != 0 || put != bytes) return HB_ERR_ARG;
        hb_hdb_volume &v = vols[(size_t)k];
        std::memset(&v, 0, sizeof(v));
        v.name = names[(size_t)k].c_str();
        v.name_len = (uint32_t)names[(size_t)k].size();
        v.bytes = bytes;
        v.volume_seqs = vs;
        v.gpu_seqs = gs;
        v.store_limit = first + gs;
        v.break_limit = first + vs;
        v.n_groups = ng;
        v.group_len = groups[(size_t)k].data();
        v.cpu_rest_bytes = o[(size_t)vs] - o[(size_t)gs];
        first += vs;
    }
    hb_hdb_index idx;
    std::memset(&idx, 0, sizeof(idx));
    idx.blocksx = blocksx;
    idx.threadsx = threadsx;
    idx.percentage = 99;
    idx.n_volumes = n_volumes;
    uint64_t len = 0;
    int rc = hb_hdbi_format(&idx, vols.data(), nullptr, 0, &len);
    if (rc) return rc;
    std::string text(len, '\0');
    rc = hb_hdbi_format(&idx, vols.data(), &text[0], len, &len);
    if (rc) return rc;
    FILE *f = std::fopen((pre + ".hDbI").c_str(), "wb");
    if (!f) return HB_ERR_ARG;
    const size_t put = std::fwrite(text.data(), 1, len, f);
    if (std::fclose(f) != 0 || put != len) return HB_ERR_ARG;
    return HB_OK;
}
