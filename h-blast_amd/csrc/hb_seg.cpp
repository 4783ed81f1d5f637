// SEG low-complexity masking of protein queries (host code, no HIP).
//
// NCBI BLAST+ masks a BLASTX query's translated frames with SEG by default
// (-seg 12 2.2 2.5, ncbi_blast_files/blast_args.cpp:309-316) before the
// lookup table is built; the stage then seeds no word inside a masked
// segment.  NCBI's implementation (algo/blast/core/blast_seg.c) is not part
// of the reference, so this restates the published algorithm (Wootton and
// Federhen, Comput. Chem. 17:149-163, 1993; Methods Enzymol. 266:554-571,
// 1996) with NCBI's protein parameters:
//   1. every window of W residues gets the Shannon entropy (bits) of its
//      composition over the 20 standard amino acids; windows with more than
//      maxbogus other letters get none;
//   2. a window at or below locut triggers a segment, which grows left and
//      right over windows at or below hicut;
//   3. the segment is trimmed to its sub-window of lowest composition
//      probability ln P = ln(assignments) + ln(permutations) - n ln 20,
//      trying at most maxtrim lengths below the full one;
//   4. the part left of the trimmed segment that still held triggers is
//      searched again on its own; the scan resumes after the segment.
// Segments never cross context boundaries; the mask is their union.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hb_seed.h"

namespace {

constexpr int kAlpha = 20;

// ncbistdaa code -> index among ACDEFGHIKLMNPQRSTVWY, -1 for any other letter
int aa_index(uint8_t c)
{
    static const int8_t map[28] = {-1, 0,  -1, 1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                                   12, 13, 14, 15, 16, 17, 18, -1, 19, -1, -1, -1, -1, -1};
    return c < 28 ? map[c] : -1;
}

struct Seg {
    const hb_seg_params &p;
    std::vector<double> lnfac;   // ln k!, k = 0..max window
    std::vector<int32_t> *out;   // segments as [begin, end] pairs (inclusive)

    // composition of seq[b, b+n): counts per letter, letters outside the 20
    struct Comp {
        int cnt[kAlpha];
        int bogus;
    };
    static void count(const uint8_t *s, int32_t b, int32_t n, Comp &c)
    {
        std::memset(&c, 0, sizeof c);
        for (int32_t i = b; i < b + n; ++i) {
            const int a = aa_index(s[i]);
            if (a < 0) ++c.bogus;
            else ++c.cnt[a];
        }
    }
    static void add(Comp &c, uint8_t r, int d)
    {
        const int a = aa_index(r);
        if (a < 0) c.bogus += d;
        else c.cnt[a] += d;
    }
    // nonzero counts, descending
    static int sorted(const Comp &c, int *sv)
    {
        int k = 0;
        for (int i = 0; i < kAlpha; ++i)
            if (c.cnt[i]) sv[k++] = c.cnt[i];
        std::sort(sv, sv + k, [](int x, int y) { return x > y; });
        return k;
    }

    static double entropy(const Comp &c)
    {
        int sv[kAlpha];
        const int k = sorted(c, sv);
        int total = 0;
        for (int i = 0; i < k; ++i) total += sv[i];
        if (total == 0) return 0.0;
        double e = 0.0;
        for (int i = 0; i < k; ++i) e += (double)sv[i] * std::log((double)sv[i] / total) / M_LN2;
        return std::fabs(e / total);
    }

    // ln of the composition's probability over a window of n letters
    double lnprob(const Comp &c, int n) const
    {
        int sv[kAlpha];
        const int k = sorted(c, sv);
        // assignments of the 20 letters to the sorted counts: 20! over the
        // factorials of the sizes of the classes of equal counts (zero too)
        double ass = lnfac[kAlpha];
        int left = kAlpha;
        for (int i = 0; i < k;) {
            int j = i;
            while (j < k && sv[j] == sv[i]) ++j;
            ass -= lnfac[j - i];
            left -= j - i;
            i = j;
        }
        ass -= lnfac[left];
        double perm = lnfac[n];
        for (int i = 0; i < k; ++i) perm -= lnfac[sv[i]];
        return ass + perm - (double)n * std::log((double)kAlpha);
    }

    // the sub-window of seq[lo, hi] with the lowest probability
    void trim(const uint8_t *s, int32_t &lo, int32_t &hi) const
    {
        const int32_t L = hi - lo + 1;
        int32_t bl = 0, br = L - 1;
        const int32_t minlen = std::max<int32_t>(1, L - p.maxtrim);
        double best = 1.0;
        for (int32_t len = L; len > minlen; --len) {
            Comp c;
            count(s, lo, len, c);
            for (int32_t i = 0;; ++i) {
                const double pr = lnprob(c, len);
                if (pr < best) {
                    best = pr;
                    bl = i;
                    br = i + len - 1;
                }
                if (i + len >= L) break;
                add(c, s[lo + i], -1);
                add(c, s[lo + i + len], +1);
            }
        }
        hi = lo + br;
        lo = lo + bl;
    }

    // segments of s[0, n) (offset: position of s[0] in the context)
    void segs(const uint8_t *s, int32_t n, int32_t offset)
    {
        const int32_t W = p.window;
        if (n < W) return;
        const int32_t down = (W + 1) / 2 - 1, up = W - down;
        const int32_t first = down, last = n - up;
        // H[i]: entropy of the window s[i - down, i - down + W), -1 if none
        std::vector<double> H((size_t)n, -1.0);
        Comp c;
        count(s, 0, W, c);
        for (int32_t i = first; i <= last; ++i) {
            if (c.bogus <= p.maxbogus) H[(size_t)i] = entropy(c);
            if (i < last) {
                add(c, s[i - down], -1);
                add(c, s[i - down + W], +1);
            }
        }
        int32_t lowlim = first;
        for (int32_t i = first; i <= last; ++i) {
            if (!(H[(size_t)i] <= p.locut && H[(size_t)i] != -1.0)) continue;
            int32_t lo = i;
            while (lo >= lowlim && H[(size_t)lo] != -1.0 && H[(size_t)lo] <= p.hicut) --lo;
            ++lo;
            int32_t hi = i;
            while (hi <= last && H[(size_t)hi] != -1.0 && H[(size_t)hi] <= p.hicut) ++hi;
            --hi;
            int32_t left = lo - down, right = hi + up - 1;
            trim(s, left, right);
            // the trigger window fell left of the trimmed segment: search
            // the part before it on its own
            if (i + up - 1 < left) {
                const int32_t l2 = lo - down, r2 = left - 1;
                segs(s + l2, r2 - l2 + 1, offset + l2);
            }
            out->push_back(left + offset);
            out->push_back(right + offset);
            i = std::min(hi, right + down);
            lowlim = i + 1;
        }
    }
};

}  // namespace

extern "C" void hb_seg_params_default(hb_seg_params *p)
{
    if (!p) return;
    p->window = 12;
    p->locut = 2.2;
    p->hicut = 2.5;
    p->maxbogus = 2;
    p->maxtrim = 100;
}

extern "C" int hb_seg_mask(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                           const int32_t *ctx_len, const hb_seg_params *params, uint8_t *mask)
{
    hb_seg_params dp;
    hb_seg_params_default(&dp);
    const hb_seg_params &p = params ? *params : dp;
    if (len < 0 || (len > 0 && (!seq || !mask)) || n_ctx < 1 || !ctx_off || !ctx_len ||
        p.window < 1 || p.window > 4096 || p.maxbogus < 0 || p.maxtrim < 0 ||
        !(p.locut <= p.hicut))
        return HB_ERR_ARG;
    for (int32_t c = 0; c < n_ctx; ++c)
        if (ctx_off[c] < 0 || ctx_len[c] < 0 || (int64_t)ctx_off[c] + ctx_len[c] > len)
            return HB_ERR_ARG;
    for (int32_t i = 0; i < len; ++i)
        if (seq[i] >= 28) return HB_ERR_RESIDUE;
    if (len) std::memset(mask, 0, (size_t)len);
    Seg S{p, {}, nullptr};
    int32_t longest = std::max(kAlpha, p.window);
    for (int32_t c = 0; c < n_ctx; ++c) longest = std::max(longest, ctx_len[c]);
    S.lnfac.assign((size_t)longest + 1, 0.0);
    for (size_t k = 1; k < S.lnfac.size(); ++k) S.lnfac[k] = S.lnfac[k - 1] + std::log((double)k);
    std::vector<int32_t> segs;
    S.out = &segs;
    for (int32_t c = 0; c < n_ctx; ++c) {
        segs.clear();
        S.segs(seq + ctx_off[c], ctx_len[c], 0);
        for (size_t k = 0; k < segs.size(); k += 2)
            for (int32_t i = segs[k]; i <= segs[k + 1]; ++i) mask[ctx_off[c] + i] = 1;
    }
    return HB_OK;
}
