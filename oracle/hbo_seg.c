/*
 * hbo_seg.c -- CPU oracle for SEG low-complexity masking of protein queries
 * (test infrastructure only; see hbo.h).
 *
 * NCBI BLAST+ masks BLASTX queries with SEG by default (-seg 12 2.2 2.5,
 * ncbi_blast_files/blast_args.cpp:309-316).  NCBI's blast_seg.c is not part
 * of the reference, so this restates the published algorithm (Wootton and
 * Federhen, Comput. Chem. 17:149-163, 1993; Methods Enzymol. 266:554-571,
 * 1996) with NCBI's protein parameters, written for clarity rather than
 * speed: every window's composition is counted from scratch, and the trim
 * tries every (length, start) pair in order.  The entropy and probability
 * formulas are exported so that tests can check them against closed forms.
 */
#include "hbo.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ACDEFGHIKLMNPQRSTVWY in ncbistdaa; every other code is "bogus" */
static int std20(uint8_t c)
{
    switch (c) {
    case 1: case 3: case 4: case 5: case 6: case 7: case 8: case 9: case 10: case 11:
    case 12: case 13: case 14: case 15: case 16: case 17: case 18: case 19: case 20: case 22:
        return 1;
    default:
        return 0;
    }
}

/* counts[28] of the letters of s[0, n); returns the number of bogus letters */
static int compose(const uint8_t *s, int32_t n, int32_t *counts)
{
    memset(counts, 0, 28 * sizeof(int32_t));
    int bogus = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (std20(s[i])) counts[s[i]]++;
        else bogus++;
    }
    return bogus;
}

/* nonzero counts in descending order */
static int desc(const int32_t *counts, int32_t *out)
{
    int k = 0;
    for (int c = 0; c < 28; ++c)
        if (counts[c]) {
            int j = k++;
            while (j > 0 && out[j - 1] < counts[c]) { out[j] = out[j - 1]; --j; }
            out[j] = counts[c];
        }
    return k;
}

static const double kLn2 = 0.69314718055994530942;   /* ln 2 */

static double lnfact(int32_t k)
{
    double r = 0;
    for (int32_t i = 2; i <= k; ++i) r += log((double)i);
    return r;
}

double hbo_seg_entropy(const int32_t *counts)
{
    int32_t sv[28];
    const int k = desc(counts, sv);
    int32_t total = 0;
    for (int i = 0; i < k; ++i) total += sv[i];
    if (total == 0) return 0;
    double e = 0;
    for (int i = 0; i < k; ++i) e += (double)sv[i] * log((double)sv[i] / total) / kLn2;
    return fabs(e / total);
}

double hbo_seg_lnprob(const int32_t *counts, int32_t n)
{
    int32_t sv[28];
    const int k = desc(counts, sv);
    /* ways to give the 20 letters these counts: 20! / prod (letters per count)! */
    double ass = lnfact(20);
    int32_t used = 0;
    int i = 0;
    while (i < k) {
        int j = i + 1;
        while (j < k && sv[j] == sv[i]) ++j;
        ass -= lnfact(j - i);
        used += j - i;
        i = j;
    }
    ass -= lnfact(20 - used);
    double perm = lnfact(n);
    for (i = 0; i < k; ++i) perm -= lnfact(sv[i]);
    return ass + perm - (double)n * log(20.0);
}

typedef struct {
    int32_t W, maxbogus, maxtrim;
    double lo, hi;
    uint8_t *mask;
} seg_run;

static void find(const seg_run *R, const uint8_t *s, int32_t n, int32_t at)
{
    const int32_t W = R->W;
    if (n < W) return;
    const int32_t down = (W + 1) / 2 - 1, up = W - down, first = down, last = n - up;
    double *H = (double *)malloc(sizeof(double) * (size_t)n);
    int32_t counts[28];
    for (int32_t i = 0; i < n; ++i) H[i] = -1;
    for (int32_t i = first; i <= last; ++i)
        if (compose(s + i - down, W, counts) <= R->maxbogus) H[i] = hbo_seg_entropy(counts);
    int32_t floor_ = first;
    for (int32_t i = first; i <= last; ++i) {
        if (H[i] == -1 || H[i] > R->lo) continue;
        int32_t a = i, b = i;
        while (a - 1 >= floor_ && H[a - 1] != -1 && H[a - 1] <= R->hi) --a;
        while (b + 1 <= last && H[b + 1] != -1 && H[b + 1] <= R->hi) ++b;
        int32_t left = a - down, right = b + up - 1;
        /* trim: lowest probability over lengths L .. minlen + 1, starts
         * ascending; the first minimum wins */
        const int32_t L = right - left + 1;
        const int32_t minlen = L - R->maxtrim > 1 ? L - R->maxtrim : 1;
        double best = 1;
        int32_t bs = 0, bl = L;
        for (int32_t len = L; len > minlen; --len)
            for (int32_t st = 0; st + len <= L; ++st) {
                compose(s + left + st, len, counts);
                const double p = hbo_seg_lnprob(counts, len);
                if (p < best) { best = p; bs = st; bl = len; }
            }
        const int32_t tl = left + bs, tr = left + bs + bl - 1;
        if (i + up - 1 < tl) find(R, s + (a - down), tl - (a - down), at + (a - down));
        for (int32_t k = tl; k <= tr; ++k) R->mask[at + k] = 1;
        i = b < tr + down ? b : tr + down;
        floor_ = i + 1;
    }
    free(H);
}

int hbo_seg_mask(const uint8_t *seq, int32_t len, int32_t n_ctx, const int32_t *ctx_off,
                 const int32_t *ctx_len, int32_t window, double locut, double hicut,
                 int32_t maxbogus, int32_t maxtrim, uint8_t *mask)
{
    if (len < 0 || n_ctx < 1 || window < 1) return HBO_ERR_ARG;
    memset(mask, 0, (size_t)len);
    seg_run R = {window, maxbogus, maxtrim, locut, hicut, mask};
    for (int32_t c = 0; c < n_ctx; ++c) {
        if (ctx_off[c] < 0 || ctx_len[c] < 0 || ctx_off[c] + ctx_len[c] > len) return HBO_ERR_ARG;
        find(&R, seq + ctx_off[c], ctx_len[c], ctx_off[c]);
    }
    return 0;
}
