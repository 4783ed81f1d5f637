// hb_synth.cpp -- deterministic synthetic protein databases and queries.
//
// Workloads of the benchmark and tests (no datasets are available offline):
//   * residues i.i.d. from the Robinson & Robinson (PNAS 1991) background
//     amino-acid frequencies;
//   * subject lengths round(exp(N(ln median, sigma))) clipped to [lo, hi];
//   * a fraction of subjects carries a planted homolog of a query segment
//     (length U[40,120], identity U[0.40,0.80], substitutions drawn from
//     P(b|a) ~ p_b exp(lambda s_ab), b != a, lambda = 0.3176, plus U{0..3}
//     indels of length U{1..3}) at a uniform position;
//   * a BLASTX query of uniform ACGT with planted coding regions.
// Every subject i is generated from its own stream seeded by (seed, i), so
// generation is parallel and any subject can be regenerated alone.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

inline uint64_t splitmix64(uint64_t &x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

struct Rng {  // xoshiro256**
    uint64_t s[4];
    explicit Rng(uint64_t seed, uint64_t stream = 0)
    {
        uint64_t x = seed ^ (stream * 0xd1342543de82ef95ull + 0x632be59bd9b4e019ull);
        for (auto &v : s) v = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t v, int k) { return (v << k) | (v >> (64 - k)); }
    inline uint64_t next()
    {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    inline double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    inline int64_t range(int64_t lo, int64_t hi)  // inclusive
    {
        return lo + (int64_t)(uniform() * (double)(hi - lo + 1));
    }
    inline double normal()
    {
        double u1 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// ncbistdaa codes of the 20 standard residues and their background
// frequencies (per 1000).
const uint8_t kCodes[20] = {1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22};
//                           A  C  D  E  F  G  H  I  K   L   M   N   P   Q   R   S   T   V   W   Y
const double kFreq[20] = {78.05, 19.25, 53.64, 62.95, 38.56, 73.77, 21.99, 51.42, 57.44, 90.19,
                          22.43, 44.87, 52.03, 42.64, 51.29, 71.20, 58.41, 64.41, 13.30, 32.16};

// BLOSUM62 scores among the 20 standard residues in kCodes order.
const int8_t kB62[20][20] = {
    // A   C   D   E   F   G   H   I   K   L   M   N   P   Q   R   S   T   V   W   Y
    {4, 0, -2, -1, -2, 0, -2, -1, -1, -1, -1, -2, -1, -1, -1, 1, 0, 0, -3, -2},   // A
    {0, 9, -3, -4, -2, -3, -3, -1, -3, -1, -1, -3, -3, -3, -3, -1, -1, -1, -2, -2},  // C
    {-2, -3, 6, 2, -3, -1, -1, -3, -1, -4, -3, 1, -1, 0, -2, 0, -1, -3, -4, -3},  // D
    {-1, -4, 2, 5, -3, -2, 0, -3, 1, -3, -2, 0, -1, 2, 0, 0, -1, -2, -3, -2},     // E
    {-2, -2, -3, -3, 6, -3, -1, 0, -3, 0, 0, -3, -4, -3, -3, -2, -2, -1, 1, 3},   // F
    {0, -3, -1, -2, -3, 6, -2, -4, -2, -4, -3, 0, -2, -2, -2, 0, -2, -3, -2, -3}, // G
    {-2, -3, -1, 0, -1, -2, 8, -3, -1, -3, -2, 1, -2, 0, 0, -1, -2, -3, -2, 2},   // H
    {-1, -1, -3, -3, 0, -4, -3, 4, -3, 2, 1, -3, -3, -3, -3, -2, -1, 3, -3, -1},  // I
    {-1, -3, -1, 1, -3, -2, -1, -3, 5, -2, -1, 0, -1, 1, 2, 0, -1, -2, -3, -2},   // K
    {-1, -1, -4, -3, 0, -4, -3, 2, -2, 4, 2, -3, -3, -2, -2, -2, -1, 1, -2, -1},  // L
    {-1, -1, -3, -2, 0, -3, -2, 1, -1, 2, 5, -2, -2, 0, -1, -1, -1, 1, -1, -1},   // M
    {-2, -3, 1, 0, -3, 0, 1, -3, 0, -3, -2, 6, -2, 0, 0, 1, 0, -3, -4, -2},       // N
    {-1, -3, -1, -1, -4, -2, -2, -3, -1, -3, -2, -2, 7, -1, -2, -1, -1, -2, -4, -3}, // P
    {-1, -3, 0, 2, -3, -2, 0, -3, 1, -2, 0, 0, -1, 5, 1, 0, -1, -2, -2, -1},      // Q
    {-1, -3, -2, 0, -3, -2, 0, -3, 2, -2, -1, 0, -2, 1, 5, -1, -1, -3, -3, -2},   // R
    {1, -1, 0, 0, -2, 0, -1, -2, 0, -2, -1, 1, -1, 0, -1, 4, 1, -2, -3, -2},      // S
    {0, -1, -1, -1, -2, -2, -2, -1, -1, -1, -1, 0, -1, -1, -1, 1, 5, 0, -2, -2},  // T
    {0, -1, -3, -2, -1, -3, -3, 3, -2, 1, 1, -3, -2, -2, -3, -2, 0, 4, -3, -1},   // V
    {-3, -2, -4, -3, 1, -2, -2, -3, -3, -2, -1, -4, -4, -2, -3, -3, -2, -3, 11, 2}, // W
    {-2, -2, -3, -2, 3, -3, 2, -1, -2, -1, -1, -2, -3, -1, -2, -2, -2, -1, 2, 7},  // Y
};

struct Tables {
    uint8_t bg[4096];        // top 12 bits of a uniform -> residue code
    double sub_cdf[20][20];  // P(b | a), b != a, cumulative
    int8_t idx_of_code[32];
    Tables()
    {
        double tot = 0;
        for (double f : kFreq) tot += f;
        double run = 0;
        int k = 0;
        for (int j = 0; j < 4096; ++j) {
            const double u = (j + 0.5) / 4096.0;
            while (k < 19 && u > (run + kFreq[k]) / tot) {
                run += kFreq[k];
                ++k;
            }
            bg[j] = kCodes[k];
        }
        for (int a = 0; a < 20; ++a) {
            double z = 0;
            for (int b = 0; b < 20; ++b)
                if (b != a) z += kFreq[b] * std::exp(0.3176 * kB62[a][b]);
            double c = 0;
            for (int b = 0; b < 20; ++b) {
                if (b != a) c += kFreq[b] * std::exp(0.3176 * kB62[a][b]) / z;
                sub_cdf[a][b] = c;
            }
        }
        for (auto &x : idx_of_code) x = -1;
        for (int i = 0; i < 20; ++i) idx_of_code[kCodes[i]] = (int8_t)i;
    }
};

const Tables &tables()
{
    static const Tables t;
    return t;
}

inline void fill_background(Rng &r, uint8_t *out, int64_t n)
{
    const Tables &T = tables();
    int64_t i = 0;
    while (i < n) {
        uint64_t x = r.next();
        for (int k = 0; k < 5 && i < n; ++k, ++i) {
            out[i] = T.bg[x >> 52];
            x <<= 12;
        }
    }
}

inline uint8_t substitute(Rng &r, uint8_t code)
{
    const Tables &T = tables();
    const int a = T.idx_of_code[code & 31];
    if (a < 0) return code;
    const double u = r.uniform();
    for (int b = 0; b < 20; ++b)
        if (b != a && u <= T.sub_cdf[a][b]) return kCodes[b];
    return kCodes[a == 19 ? 18 : 19];
}

// Homolog of src[0..len): identity fraction kept, the rest substituted, then
// n_indel indels of length 1..3.  Returns the new length (<= cap).
int32_t make_homolog(Rng &r, const uint8_t *src, int32_t len, double identity, int n_indel,
                     uint8_t *out, int32_t cap)
{
    std::vector<uint8_t> v(src, src + len);
    for (auto &c : v)
        if (r.uniform() >= identity) c = substitute(r, c);
    for (int k = 0; k < n_indel && !v.empty(); ++k) {
        const int32_t L = (int32_t)r.range(1, 3);
        const int32_t p = (int32_t)r.range(0, (int64_t)v.size() - 1);
        if (r.uniform() < 0.5) {
            std::vector<uint8_t> ins((size_t)L);
            fill_background(r, ins.data(), L);
            v.insert(v.begin() + p, ins.begin(), ins.end());
        } else {
            const int32_t e = std::min<int32_t>((int32_t)v.size(), p + L);
            v.erase(v.begin() + p, v.begin() + e);
        }
    }
    const int32_t n = std::min<int32_t>((int32_t)v.size(), cap);
    memcpy(out, v.data(), (size_t)n);
    return n;
}

void gen_subject(uint64_t seed, uint32_t i, uint8_t *out, int64_t L, const uint8_t *pq,
                 int32_t pqlen, double frac)
{
    Rng r(seed, (uint64_t)i + 1);
    fill_background(r, out, L);
    if (!pq || pqlen < 40 || frac <= 0 || r.uniform() >= frac) return;
    const int32_t seg = (int32_t)std::min<int64_t>(r.range(40, 120), pqlen);
    const int32_t qs = (int32_t)r.range(0, pqlen - seg);
    const double ident = 0.40 + 0.40 * r.uniform();
    const int n_indel = (int)r.range(0, 3);
    uint8_t buf[256];
    int32_t h = make_homolog(r, pq + qs, seg, ident, n_indel, buf, (int32_t)sizeof(buf));
    if (h > L) h = (int32_t)L;
    if (h <= 0) return;
    const int64_t pos = r.range(0, L - h);
    memcpy(out + pos, buf, (size_t)h);
}

}  // namespace

extern "C" {

// Log-normal lengths round(exp(N(ln median, sigma))) clipped to [lo, hi].
// Returns the sum of the lengths.
uint64_t hbs_lognormal_lengths(uint64_t seed, uint32_t n, double median, double sigma,
                               uint32_t lo, uint32_t hi, uint32_t *out)
{
    uint64_t tot = 0;
    const double mu = std::log(median);
    for (uint32_t i = 0; i < n; ++i) {
        Rng r(seed ^ 0x6c656e67746873ull, i);
        double x = std::round(std::exp(mu + sigma * r.normal()));
        if (x < lo) x = lo;
        if (x > hi) x = hi;
        out[i] = (uint32_t)x;
        tot += out[i];
    }
    return tot;
}

// Residues i.i.d. from the background frequencies.
void hbs_random_protein(uint64_t seed, uint32_t len, uint8_t *out)
{
    Rng r(seed, 0);
    fill_background(r, out, len);
}

// Fills subjects [0, n) of a database laid out by offsets (n+1 entries).
void hbs_db_fill(uint64_t seed, uint32_t n, const uint64_t *offsets, uint8_t *residues,
                 const uint8_t *plant_query, int32_t plant_qlen, double plant_frac, int n_threads)
{
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 64) n_threads = 64;
    auto work = [&](uint32_t t) {
        for (uint32_t i = t; i < n; i += (uint32_t)n_threads)
            gen_subject(seed, i, residues + offsets[i], (int64_t)(offsets[i + 1] - offsets[i]),
                        plant_query, plant_qlen, plant_frac);
    };
    if (n_threads == 1) {
        work(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < n_threads; ++t) th.emplace_back(work, (uint32_t)t);
    for (auto &x : th) x.join();
}

// Homolog of a segment (identity, indel count given); returns its length.
int32_t hbs_homolog(uint64_t seed, const uint8_t *src, int32_t len, double identity, int n_indel,
                    uint8_t *out, int32_t cap)
{
    Rng r(seed, 7);
    return make_homolog(r, src, len, identity, n_indel, out, cap);
}

// BLASTX query: nt_len uniform ACGT with the coding regions of two protein
// segments planted, the first on the plus strand and the second reverse
// complemented (each codon drawn uniformly among its synonyms).
// Returns 0, or -1 if the regions do not fit.
int hbs_blastx_query(uint64_t seed, int32_t nt_len, const uint8_t *seg1, int32_t l1,
                     const uint8_t *seg2, int32_t l2, char *out)
{
    static const char kStdaa[] = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";
    static const char kCode1[] =
        "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";
    static const char kNt[] = "TCAG";
    Rng r(seed, 11);
    for (int32_t i = 0; i < nt_len; ++i) out[i] = "ACGT"[r.next() >> 62];
    const int32_t n1 = 3 * l1, n2 = 3 * l2;
    if (n1 + n2 > nt_len) return -1;
    auto reverse_translate = [&](const uint8_t *seg, int32_t l, std::vector<char> &nt) {
        nt.clear();
        for (int32_t k = 0; k < l; ++k) {
            const char aa = kStdaa[seg[k] & 31];
            int cand[64], nc = 0;
            for (int c = 0; c < 64; ++c)
                if (kCode1[c] == aa) cand[nc++] = c;
            const int c = nc ? cand[r.range(0, nc - 1)] : 0;
            nt.push_back(kNt[c >> 4]);
            nt.push_back(kNt[(c >> 2) & 3]);
            nt.push_back(kNt[c & 3]);
        }
    };
    std::vector<char> a, b;
    reverse_translate(seg1, l1, a);
    reverse_translate(seg2, l2, b);
    // region 1 in the first part of the query, region 2 after it
    const int32_t p1 = (int32_t)r.range(0, nt_len - n1 - n2);
    const int32_t p2 = (int32_t)r.range(p1 + n1, nt_len - n2);
    memcpy(out + p1, a.data(), (size_t)n1);
    for (int32_t k = 0; k < n2; ++k) {
        const char ch = b[n2 - 1 - k];
        out[p2 + k] = ch == 'A' ? 'T' : ch == 'T' ? 'A' : ch == 'C' ? 'G' : 'C';
    }
    return 0;
}

}  // extern "C"
