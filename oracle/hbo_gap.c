/*
 * hbo_gap.c -- CPU oracle for the preliminary gapped extension (test
 * infrastructure only; see hbo.h).
 *
 * Per subject, the in-frame protein path of BLAST_GetGappedScore
 * (ncbi_blast_files/blast_gapalign.c:3233-3560, called per subject by
 * ncbi_blast_files/blast_engine.c:617 on the HSPs the seed stage saved at
 * :562-584), score only, without restricted (approximate) alignment:
 *
 *   start point          BlastGetStartForGappedAlignment      blast_gapalign.c:3184-3229
 *   subject window       AdjustSubjectRange                   blast_gapalign.c:3608-3637
 *   two-sided extension  s_BlastProtGappedAlignment (in-frame) blast_gapalign.c:3654-3783
 *   X-drop DP            Blast_SemiGappedAlign, score only    blast_gapalign.c:711-937
 *   context offsets      s_AdjustHspOffsetsAndGetQueryData    blast_gapalign.c:2420-2440
 *   per-HSP loop         BLAST_GetGappedScore                 blast_gapalign.c:3362-3549
 *
 * Two steps live in NCBI BLAST+ 2.2.28 files that the reference does not
 * vendor, and are restated from that release's published source:
 *   - Blast_InitHitListSortByScore (blast_extend.c), applied by the caller
 *     at blast_engine.c:584: ungapped HSPs ordered by score descending, then
 *     subject start ascending, length descending, query start ascending;
 *   - BlastIntervalTreeContainsHSP (blast_itree.c) with
 *     min_diag_separation 0: an ungapped HSP is skipped if a saved gapped HSP
 *     of the same context scores at least as much and its query x subject box
 *     holds both the HSP's start point (q_start, s_start) and its end point
 *     (q_start + len, s_start + len), ends inclusive.  The tree only speeds
 *     up that search; a linear scan over the saved HSPs gives the same answer.
 */
#include "hbo.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NEG_INF (INT32_MIN / 2) /* MININT, blast_gapalign.c:61 */
#define START_WINDOW 11         /* HSP_MAX_WINDOW of NCBI blast_gapalign.h */
#define LONG_SUBJECT 90000      /* MAX_SUBJECT_OFFSET, blast_gapalign.c:3600 */
#define GAP_ALLOWANCE 3000      /* MAX_TOTAL_GAPS, blast_gapalign.c:3605 */

static const int8_t *mat_of(const hbo_query *q) { return q->matrix ? q->matrix : hbo_blosum62(); }

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

int32_t hbo_semigapped(const int8_t *m, const uint8_t *A, const uint8_t *B, int32_t M,
                       int32_t N, int reverse, int32_t gap_open, int32_t gap_extend,
                       int32_t x_dropoff, int32_t *a_off, int32_t *b_off)
{
    const int32_t goe = gap_open + gap_extend;
    if (x_dropoff < goe) x_dropoff = goe;
    *a_off = 0;
    *b_off = 0;
    if (N <= 0 || M <= 0) return 0;
    /* column c (0..N) holds the row's best score ending at B letter c and the
       best score of an alignment ending in a gap in A at that column */
    int32_t *best = (int32_t *)malloc(sizeof(int32_t) * ((size_t)N + 2));
    int32_t *vgap = (int32_t *)malloc(sizeof(int32_t) * ((size_t)N + 2));
    if (!best || !vgap) {
        free(best);
        free(vgap);
        return NEG_INF;
    }
    /* row 0: leading gaps in A */
    best[0] = 0;
    vgap[0] = -goe;
    int32_t end = 1, v = -goe;
    for (; end <= N; ++end) {
        if (v < -x_dropoff) break;
        best[end] = v;
        vgap[end] = v - goe;
        v -= gap_extend;
    }
    int32_t lo = 0, top = 0;
    for (int32_t a = 1; a <= M; ++a) {
        const int8_t *row = m + HBO_ALPHA * (reverse ? A[M - a] : A[a]);
        int32_t diag = NEG_INF, hgap = NEG_INF, last = lo;
        for (int32_t b = lo; b < end; ++b) {
            /* diagonal input of column b + 1 (no column N + 1 exists) */
            int32_t next = NEG_INF;
            if (b < N) next = best[b] + row[reverse ? B[N - 1 - b] : B[b + 1]];
            int32_t col = vgap[b];
            int32_t s = diag;
            if (s < col) s = col;
            if (s < hgap) s = hgap;
            if (top - s > x_dropoff) {
                /* pruned: a leading run moves the band's left edge */
                if (b == lo)
                    ++lo;
                else
                    best[b] = NEG_INF;
            } else {
                last = b;
                if (s > top) {
                    top = s;
                    *a_off = a;
                    *b_off = b;
                }
                hgap -= gap_extend;
                col -= gap_extend;
                vgap[b] = imax(s - goe, col);
                hgap = imax(s - goe, hgap);
                best[b] = s;
            }
            diag = next;
        }
        if (lo == end) break;
        if (last < end - 1) {
            end = last + 1;
        } else {
            /* the row's last cell survived: open the columns to its right
               that the running gap in B keeps within the drop-off */
            while (hgap >= top - x_dropoff && end <= N) {
                best[end] = hgap;
                vgap[end] = hgap - goe;
                hgap -= gap_extend;
                ++end;
            }
        }
        if (end <= N) {
            best[end] = NEG_INF;
            vgap[end] = NEG_INF;
            ++end;
        }
    }
    free(best);
    free(vgap);
    return top;
}

int32_t hbo_gapped_start(const int8_t *m, const uint8_t *q, const uint8_t *s, int32_t q_start,
                         int32_t q_len, int32_t s_start, int32_t s_len)
{
    if (q_len <= START_WINDOW) return q_start + q_len / 2;
    int32_t sc = 0;
    for (int32_t i = 0; i < START_WINDOW; ++i)
        sc += m[HBO_ALPHA * q[q_start + i] + s[s_start + i]];
    int32_t top = sc, at = q_start + START_WINDOW - 1;
    const int32_t n = imin(q_len, s_len);
    for (int32_t i = START_WINDOW; i < n; ++i) {
        sc -= m[HBO_ALPHA * q[q_start + i - START_WINDOW] + s[s_start + i - START_WINDOW]];
        sc += m[HBO_ALPHA * q[q_start + i] + s[s_start + i]];
        if (sc > top) {
            top = sc;
            at = q_start + i;
        }
    }
    return top > 0 ? at - START_WINDOW / 2 : q_start;
}

/* AdjustSubjectRange: for very long subjects, restrict the extension to the
   part of the subject the query can reach with a bounded number of gaps. */
static void subject_window(int32_t *s_len_left, int32_t *subject_len, int32_t q_len_left,
                           int32_t q_len, int32_t *shift)
{
    const int32_t slen = *subject_len;
    *shift = 0;
    if (slen < LONG_SUBJECT) return;
    const int32_t s_off = *s_len_left;
    const int32_t reach_left = q_len_left + GAP_ALLOWANCE;
    const int32_t reach_right = q_len - q_len_left + GAP_ALLOWANCE;
    if (s_off > reach_left) {
        *shift = s_off - reach_left;
        *s_len_left = reach_left;
    }
    *subject_len = imin(slen, s_off + reach_right) - *shift;
}

int32_t hbo_prot_gapped(const int8_t *m, const uint8_t *q, int32_t q_len, const uint8_t *s,
                        int32_t s_len, int32_t q_off, int32_t s_off,
                        const hbo_gap_params *gp, int32_t *coords)
{
    int32_t ql = q_off + 1, sl = s_off + 1, sub_len = s_len, shift = 0, a, b;
    subject_window(&sl, &sub_len, ql, q_len, &shift);
    /* left: from (q_off, s_off) backwards, that pair included */
    const int32_t left = hbo_semigapped(m, q, s + shift, ql, sl, 1, gp->gap_open,
                                        gp->gap_extend, gp->x_dropoff, &a, &b);
    coords[0] = ql - a;
    coords[2] = sl - b + shift;
    int32_t right = 0;
    if (ql < q_len && sl < sub_len) {
        right = hbo_semigapped(m, q + q_off, s + s_off, q_len - ql, sub_len - sl, 0,
                               gp->gap_open, gp->gap_extend, gp->x_dropoff, &a, &b);
        coords[1] = a + q_off + 1;
        coords[3] = b + s_off + 1;
    } else {
        coords[1] = q_off;
        coords[3] = s_off;
    }
    return left + right;
}

/* Ungapped HSP order of Blast_InitHitListSortByScore. */
static int init_order(const void *pa, const void *pb)
{
    const hbo_hsp *x = (const hbo_hsp *)pa, *y = (const hbo_hsp *)pb;
    if (x->score != y->score) return x->score > y->score ? -1 : 1;
    if (x->s_start != y->s_start) return x->s_start < y->s_start ? -1 : 1;
    if (x->len != y->len) return x->len > y->len ? -1 : 1;
    if (x->q_start != y->q_start) return x->q_start < y->q_start ? -1 : 1;
    return 0;
}

static int inside(const hbo_ghsp *g, int32_t qp, int32_t sp)
{
    return g->q_start <= qp && g->q_end >= qp && g->s_start <= sp && g->s_end >= sp;
}

int hbo_gapped_subject(const hbo_query *q, const uint8_t *subj, int32_t slen, const hbo_hsp *in,
                       int32_t n_in, const hbo_gap_params *gp, const int32_t *gap_cutoff,
                       hbo_ghsp *out, int32_t *n_out, int32_t *extensions)
{
    *n_out = 0;
    *extensions = 0;
    if (n_in <= 0) return 0;
    hbo_hsp *h = (hbo_hsp *)malloc(sizeof(hbo_hsp) * (size_t)n_in);
    if (!h) return HBO_ERR_NOMEM;
    memcpy(h, in, sizeof(hbo_hsp) * (size_t)n_in);
    qsort(h, (size_t)n_in, sizeof(hbo_hsp), init_order);
    const int8_t *m = mat_of(q);
    int32_t saved = 0, ext = 0;
    for (int32_t i = 0; i < n_in; ++i) {
        const int32_t c = h[i].ctx;
        if (c < 0 || c >= q->n_ctx) {
            free(h);
            return HBO_ERR_ARG;
        }
        const int32_t base = q->ctx_off[c];
        const uint8_t *cq = q->seq + base;
        const int32_t clen = q->ctx_len[c];
        const int32_t qs = h[i].q_start - base, ss = h[i].s_start, len = h[i].len;
        int contained = 0;
        for (int32_t j = 0; j < saved && !contained; ++j) {
            const hbo_ghsp *g = &out[j];
            contained = g->ctx == c && h[i].score <= g->score && inside(g, qs, ss) &&
                        inside(g, qs + len, ss + len);
        }
        if (contained) continue;
        ++ext;
        const int32_t q_off0 = h[i].q_off - base;
        const int32_t q_off = hbo_gapped_start(m, cq, subj, qs, len, ss, len);
        const int32_t s_off = h[i].s_off + (q_off - q_off0);
        int32_t co[4];
        const int32_t score = hbo_prot_gapped(m, cq, clen, subj, slen, q_off, s_off, gp, co);
        if (score >= gap_cutoff[c]) {
            hbo_ghsp *g = &out[saved++];
            g->oid = h[i].oid;
            g->ctx = c;
            g->q_start = co[0];
            g->q_end = co[1];
            g->s_start = co[2];
            g->s_end = co[3];
            g->q_gapped_start = q_off;
            g->s_gapped_start = s_off;
            g->score = score;
            g->pad = 0;
        }
    }
    free(h);
    *n_out = saved;
    *extensions = ext;
    return 0;
}

typedef struct {
    const hbo_query *q;
    const uint8_t *res;
    const uint64_t *offs;
    const hbo_hsp *in;
    const uint64_t *first; /* per subject: first input record, n_subjects + 1 */
    uint32_t n_subjects;
    const hbo_gap_params *gp;
    const int32_t *cutoff;
    hbo_ghsp *out;          /* slot i of subject k at first[k] + i */
    int32_t *counts;        /* per subject: saved, extensions */
    int next;               /* shared subject cursor (chunks of 64) */
    pthread_mutex_t mu;
    int err;
} gap_job;

static void *gap_worker(void *arg)
{
    gap_job *J = (gap_job *)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        const uint32_t k0 = (uint32_t)J->next;
        J->next += 64;
        pthread_mutex_unlock(&J->mu);
        if (k0 >= J->n_subjects) break;
        const uint32_t k1 = k0 + 64 < J->n_subjects ? k0 + 64 : J->n_subjects;
        for (uint32_t k = k0; k < k1; ++k) {
            const uint64_t f = J->first[k], n = J->first[k + 1] - f;
            if (!n) continue;
            const uint8_t *s = J->res + J->offs[k];
            const int32_t slen = (int32_t)(J->offs[k + 1] - J->offs[k]);
            const int rc = hbo_gapped_subject(J->q, s, slen, J->in + f, (int32_t)n, J->gp,
                                              J->cutoff, J->out + f, &J->counts[2 * k],
                                              &J->counts[2 * k + 1]);
            if (rc) {
                pthread_mutex_lock(&J->mu);
                J->err = rc;
                pthread_mutex_unlock(&J->mu);
            }
        }
    }
    return NULL;
}

int64_t hbo_gapped_db(const hbo_query *q, const uint8_t *residues, const uint64_t *offsets,
                      uint32_t n_subjects, const hbo_hsp *in, uint64_t n_in,
                      const hbo_gap_params *gp, const int32_t *gap_cutoff, int n_threads,
                      hbo_ghsp **out, int32_t *counts)
{
    if (!q || !residues || !offsets || !gp || !gap_cutoff || !out || !counts ||
        (n_in && !in))
        return HBO_ERR_ARG;
    *out = NULL;
    uint64_t *first = (uint64_t *)calloc((size_t)n_subjects + 1, sizeof(uint64_t));
    hbo_ghsp *tmp = (hbo_ghsp *)malloc(sizeof(hbo_ghsp) * (size_t)(n_in ? n_in : 1));
    if (!first || !tmp) {
        free(first);
        free(tmp);
        return HBO_ERR_NOMEM;
    }
    for (uint64_t i = 0; i < n_in; ++i) {
        if (in[i].oid >= n_subjects || (i && in[i].oid < in[i - 1].oid)) {
            free(first);
            free(tmp);
            return HBO_ERR_ARG; /* records must be grouped by ascending oid */
        }
        first[in[i].oid + 1]++;
    }
    for (uint32_t k = 0; k < n_subjects; ++k) first[k + 1] += first[k];
    memset(counts, 0, sizeof(int32_t) * 2 * (size_t)n_subjects);
    gap_job J = {q, residues, offsets, in, first, n_subjects, gp, gap_cutoff, tmp, counts,
                 0, PTHREAD_MUTEX_INITIALIZER, 0};
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 1; t < n_threads; ++t)
        if (pthread_create(&th[started], NULL, gap_worker, &J) == 0) ++started;
    gap_worker(&J);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    if (J.err) {
        free(first);
        free(tmp);
        return J.err;
    }
    /* compact: subject k's saved HSPs are tmp[first[k] .. first[k] + counts[2k]) */
    uint64_t n = 0;
    for (uint32_t k = 0; k < n_subjects; ++k) {
        const int32_t c = counts[2 * k];
        if (c && n != first[k]) memmove(tmp + n, tmp + first[k], sizeof(hbo_ghsp) * (size_t)c);
        n += (uint64_t)c;
    }
    free(first);
    *out = tmp;
    return (int64_t)n;
}
