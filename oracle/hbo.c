/*
 * hbo.c -- CPU oracle (test infrastructure only; see hbo.h).
 *
 * Sequential restatement of the two-hit ungapped word finder.  The NCBI
 * BLAST+ 2.2.28 CPU word finder (aa_ungapped.c, blast_aalookup.c,
 * blast_extend.c; fetched from NCBI's FTP by h_blast_v1.1/install:27,79 and
 * absent offline) is the named oracle; each step below follows the
 * reference's own in-tree statement of that algorithm:
 *
 *   matrix               h_blast_v1.1/H-BLAST/h_blast.c:42-71
 *   3-mer index          H-BLAST/h_blast_kernel.cu:169-189, 646-649 (CHARSIZE 5, :3)
 *   cell contents        H-BLAST/h_blast.cu:940-958, 165-194 (entries ascending)
 *   two-hit diag machine H-BLAST/h_blast_kernel.cu:672-870 (k1), 218-336 (k2)
 *   context search       H-BLAST/h_blast_kernel.cu:368-380, 706-721
 *   extend two-hit       H-BLAST/h_blast_kernel.cu:1069-1130
 *   extend right         H-BLAST/h_blast_kernel.cu:1133-1326
 *   extend left          H-BLAST/h_blast_kernel.cu:1329-1540
 *   cutoff / record      H-BLAST/h_blast_kernel.cu:736-760; ncbi_blast_files/blast_engine.c:562-576
 *   diag array sizing    H-BLAST/h_blast.cu:959-966 (NCBI diag_table: pow2 >= qlen + window)
 *
 * The reference GPU kernel defers extensions to odd s_off
 * (h_blast_kernel.cu:873-942); this oracle does not: every hit is handled in
 * subject order, then query order, which is the order of the CPU word finder
 * named as the oracle.
 */
#include "hbo.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* BLOSUM62 in ncbistdaa order -,A,B,C,D,E,F,G,H,I,K,L,M,N,P,Q,R,S,T,V,W,X,Y,Z,U,*,O,J
 * (values of h_blast_v1.1/H-BLAST/h_blast.c:42-71, without its 4 padding
 * columns; the gap letter scores -128 against everything). */
static const int8_t kBlosum62[HBO_ALPHA][HBO_ALPHA] = {
    /* - */ {-128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128, -128},
    /* A */ {-128,    4,   -2,    0,   -2,   -1,   -2,    0,   -2,   -1,   -1,   -1,   -1,   -2,   -1,   -1,   -1,    1,    0,    0,   -3,   -1,   -2,   -1,   -1,   -4,   -1,   -1},
    /* B */ {-128,   -2,    4,   -3,    4,    1,   -3,   -1,    0,   -3,    0,   -4,   -3,    4,   -2,    0,   -1,    0,   -1,   -3,   -4,   -1,   -3,    0,   -1,   -4,   -1,   -3},
    /* C */ {-128,    0,   -3,    9,   -3,   -4,   -2,   -3,   -3,   -1,   -3,   -1,   -1,   -3,   -3,   -3,   -3,   -1,   -1,   -1,   -2,   -1,   -2,   -3,   -1,   -4,   -1,   -1},
    /* D */ {-128,   -2,    4,   -3,    6,    2,   -3,   -1,   -1,   -3,   -1,   -4,   -3,    1,   -1,    0,   -2,    0,   -1,   -3,   -4,   -1,   -3,    1,   -1,   -4,   -1,   -3},
    /* E */ {-128,   -1,    1,   -4,    2,    5,   -3,   -2,    0,   -3,    1,   -3,   -2,    0,   -1,    2,    0,    0,   -1,   -2,   -3,   -1,   -2,    4,   -1,   -4,   -1,   -3},
    /* F */ {-128,   -2,   -3,   -2,   -3,   -3,    6,   -3,   -1,    0,   -3,    0,    0,   -3,   -4,   -3,   -3,   -2,   -2,   -1,    1,   -1,    3,   -3,   -1,   -4,   -1,    0},
    /* G */ {-128,    0,   -1,   -3,   -1,   -2,   -3,    6,   -2,   -4,   -2,   -4,   -3,    0,   -2,   -2,   -2,    0,   -2,   -3,   -2,   -1,   -3,   -2,   -1,   -4,   -1,   -4},
    /* H */ {-128,   -2,    0,   -3,   -1,    0,   -1,   -2,    8,   -3,   -1,   -3,   -2,    1,   -2,    0,    0,   -1,   -2,   -3,   -2,   -1,    2,    0,   -1,   -4,   -1,   -3},
    /* I */ {-128,   -1,   -3,   -1,   -3,   -3,    0,   -4,   -3,    4,   -3,    2,    1,   -3,   -3,   -3,   -3,   -2,   -1,    3,   -3,   -1,   -1,   -3,   -1,   -4,   -1,    3},
    /* K */ {-128,   -1,    0,   -3,   -1,    1,   -3,   -2,   -1,   -3,    5,   -2,   -1,    0,   -1,    1,    2,    0,   -1,   -2,   -3,   -1,   -2,    1,   -1,   -4,   -1,   -3},
    /* L */ {-128,   -1,   -4,   -1,   -4,   -3,    0,   -4,   -3,    2,   -2,    4,    2,   -3,   -3,   -2,   -2,   -2,   -1,    1,   -2,   -1,   -1,   -3,   -1,   -4,   -1,    3},
    /* M */ {-128,   -1,   -3,   -1,   -3,   -2,    0,   -3,   -2,    1,   -1,    2,    5,   -2,   -2,    0,   -1,   -1,   -1,    1,   -1,   -1,   -1,   -1,   -1,   -4,   -1,    2},
    /* N */ {-128,   -2,    4,   -3,    1,    0,   -3,    0,    1,   -3,    0,   -3,   -2,    6,   -2,    0,    0,    1,    0,   -3,   -4,   -1,   -2,    0,   -1,   -4,   -1,   -3},
    /* P */ {-128,   -1,   -2,   -3,   -1,   -1,   -4,   -2,   -2,   -3,   -1,   -3,   -2,   -2,    7,   -1,   -2,   -1,   -1,   -2,   -4,   -1,   -3,   -1,   -1,   -4,   -1,   -3},
    /* Q */ {-128,   -1,    0,   -3,    0,    2,   -3,   -2,    0,   -3,    1,   -2,    0,    0,   -1,    5,    1,    0,   -1,   -2,   -2,   -1,   -1,    4,   -1,   -4,   -1,   -2},
    /* R */ {-128,   -1,   -1,   -3,   -2,    0,   -3,   -2,    0,   -3,    2,   -2,   -1,    0,   -2,    1,    5,   -1,   -1,   -3,   -3,   -1,   -2,    0,   -1,   -4,   -1,   -2},
    /* S */ {-128,    1,    0,   -1,    0,    0,   -2,    0,   -1,   -2,    0,   -2,   -1,    1,   -1,    0,   -1,    4,    1,   -2,   -3,   -1,   -2,    0,   -1,   -4,   -1,   -2},
    /* T */ {-128,    0,   -1,   -1,   -1,   -1,   -2,   -2,   -2,   -1,   -1,   -1,   -1,    0,   -1,   -1,   -1,    1,    5,    0,   -2,   -1,   -2,   -1,   -1,   -4,   -1,   -1},
    /* V */ {-128,    0,   -3,   -1,   -3,   -2,   -1,   -3,   -3,    3,   -2,    1,    1,   -3,   -2,   -2,   -3,   -2,    0,    4,   -3,   -1,   -1,   -2,   -1,   -4,   -1,    2},
    /* W */ {-128,   -3,   -4,   -2,   -4,   -3,    1,   -2,   -2,   -3,   -3,   -2,   -1,   -4,   -4,   -2,   -3,   -3,   -2,   -3,   11,   -1,    2,   -2,   -1,   -4,   -1,   -2},
    /* X */ {-128,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -4,   -1,   -1},
    /* Y */ {-128,   -2,   -3,   -2,   -3,   -2,    3,   -3,    2,   -1,   -2,   -1,   -1,   -2,   -3,   -1,   -2,   -2,   -2,   -1,    2,   -1,    7,   -2,   -1,   -4,   -1,   -1},
    /* Z */ {-128,   -1,    0,   -3,    1,    4,   -3,   -2,    0,   -3,    1,   -3,   -1,    0,   -1,    4,    0,    0,   -1,   -2,   -2,   -1,   -2,    4,   -1,   -4,   -1,   -3},
    /* U */ {-128,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -4,   -1,   -1},
    /* * */ {-128,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,   -4,    1,   -4,   -4},
    /* O */ {-128,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -1,   -4,   -1,   -1},
    /* J */ {-128,   -1,   -3,   -1,   -3,   -3,    0,   -4,   -3,    3,   -3,    3,    2,   -3,   -3,   -2,   -2,   -2,   -1,    2,   -2,   -1,   -1,   -3,   -1,   -4,   -1,    3}
};

const int8_t *hbo_blosum62(void) { return &kBlosum62[0][0]; }

static const int8_t *matrix_of(const hbo_query *q)
{
    return q->matrix ? q->matrix : &kBlosum62[0][0];
}

static int check_query(const hbo_query *q)
{
    if (!q || !q->seq || q->len <= 0 || q->n_ctx <= 0 || !q->ctx_off ||
        !q->ctx_len || !q->xdrop || !q->cutoff || q->word_size != 3 ||
        q->window <= q->word_size)
        return HBO_ERR_ARG;
    for (int32_t i = 0; i < q->len; ++i)
        if (q->seq[i] >= HBO_ALPHA) return HBO_ERR_RESIDUE;
    int32_t prev_end = 0;
    for (int32_t k = 0; k < q->n_ctx; ++k) {
        if (q->ctx_len[k] < 0 || q->ctx_off[k] < prev_end ||
            q->ctx_off[k] + q->ctx_len[k] > q->len)
            return HBO_ERR_ARG;
        prev_end = q->ctx_off[k] + q->ctx_len[k];
    }
    return 0;
}

/* Neighbourhood words: a query word at offset o (entirely inside one
 * context) is recorded in cell w when sum_i M[q[o+i]][w_i] >= T, or when w
 * equals the query word itself (an exact word always seeds, also when its
 * self-score is below T).  Query offsets are visited in ascending order so
 * each cell lists them ascending, as lut_sorting leaves them
 * (H-BLAST/h_blast.cu:165-194). */
int hbo_lut_build(const hbo_query *q, hbo_lut *out)
{
    int rc = check_query(q);
    if (rc) return rc;
    const int8_t *m = matrix_of(q);
    const int32_t T = q->threshold;
    uint32_t *count = (uint32_t *)calloc(HBO_NCELLS + 1, sizeof(uint32_t));
    if (!count) return HBO_ERR_NOMEM;
    int8_t rowmax[HBO_ALPHA];
    for (int a = 0; a < HBO_ALPHA; ++a) {
        int8_t mx = -128;
        for (int b = 0; b < HBO_ALPHA; ++b)
            if (m[a * HBO_ALPHA + b] > mx) mx = m[a * HBO_ALPHA + b];
        rowmax[a] = mx;
    }
    uint32_t *entries = NULL;
    uint32_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        uint32_t *fill = NULL;
        if (pass == 1) {
            uint32_t run = 0;
            for (int c = 0; c < HBO_NCELLS; ++c) {
                uint32_t n = count[c];
                count[c] = run;
                run += n;
            }
            count[HBO_NCELLS] = run;
            total = run;
            entries = (uint32_t *)malloc((size_t)(total ? total : 1) * sizeof(uint32_t));
            fill = (uint32_t *)malloc(HBO_NCELLS * sizeof(uint32_t));
            if (!entries || !fill) {
                free(entries); free(fill); free(count);
                return HBO_ERR_NOMEM;
            }
            memcpy(fill, count, HBO_NCELLS * sizeof(uint32_t));
        }
        for (int32_t k = 0; k < q->n_ctx; ++k) {
            int32_t lo = q->ctx_off[k], hi = q->ctx_off[k] + q->ctx_len[k] - 3;
            for (int32_t o = lo; o <= hi; ++o) {
                /* NCBI indexes words of the unmasked segments only */
                if (q->lut_mask && (q->lut_mask[o] | q->lut_mask[o + 1] | q->lut_mask[o + 2]))
                    continue;
                const uint8_t q0 = q->seq[o], q1 = q->seq[o + 1], q2 = q->seq[o + 2];
                const uint32_t self = ((uint32_t)q0 << 10) | ((uint32_t)q1 << 5) | q2;
                const int8_t *r0 = m + q0 * HBO_ALPHA, *r1 = m + q1 * HBO_ALPHA, *r2 = m + q2 * HBO_ALPHA;
                int self_seen = 0;
                for (int a = 0; a < HBO_ALPHA; ++a) {
                    int32_t sa = r0[a];
                    if (sa + rowmax[q1] + rowmax[q2] < T && a != q0) continue;
                    for (int b = 0; b < HBO_ALPHA; ++b) {
                        int32_t sb = sa + r1[b];
                        if (sb + rowmax[q2] < T && !(a == q0 && b == q1)) continue;
                        for (int c = 0; c < HBO_ALPHA; ++c) {
                            uint32_t w = ((uint32_t)a << 10) | ((uint32_t)b << 5) | (uint32_t)c;
                            if (sb + r2[c] >= T || w == self) {
                                if (w == self) self_seen = 1;
                                if (pass == 0) count[w]++;
                                else entries[fill[w]++] = (uint32_t)o;
                            }
                        }
                    }
                }
                (void)self_seen;
            }
        }
        free(fill);
    }
    out->start = count;
    out->entries = entries;
    out->n_entries = total;
    return 0;
}

void hbo_lut_free(hbo_lut *lut)
{
    if (!lut) return;
    free(lut->start);
    free(lut->entries);
    lut->start = NULL;
    lut->entries = NULL;
    lut->n_entries = 0;
}

/* Left X-drop walk from (s_off, q_off) towards offset 0
 * (H-BLAST/h_blast_kernel.cu:1329-1540): n = min(s_off, q_off), positions
 * i = n..0, strict '>' updates the best, stop once max - score >= dropoff;
 * length = n - best_i + 1 (0 if nothing improved).  At n = 0 this walks
 * position 0, as NCBI's loop (for i = n; i >= 0; i--) does; the reference
 * kernel guards its walk with if (n > 0) (h_blast_kernel.cu:1368) and
 * returns length 0 there.  The two-hit path never reaches n < 2 (the
 * trigger lies at least 3 residues after the first hit on both sequences),
 * so the difference does not show in any HSP. */
int32_t hbo_extend_left(const int8_t *m, const uint8_t *s, const uint8_t *q,
                        int32_t s_off, int32_t q_off, int32_t dropoff,
                        int32_t maxscore, int32_t *length)
{
    int32_t n = s_off < q_off ? s_off : q_off;
    int32_t best_i = n + 1;
    int32_t score = maxscore;
    const uint8_t *sp = s + s_off - n, *qp = q + q_off - n;
    for (int32_t i = n; i >= 0; --i) {
        score += m[qp[i] * HBO_ALPHA + sp[i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (maxscore - score >= dropoff) break;
    }
    *length = n - best_i + 1;
    return maxscore;
}

/* Right X-drop walk (H-BLAST/h_blast_kernel.cu:1133-1326): n = min(slen -
 * s_off, qlen - q_off), stop when score <= 0 or max - score >= dropoff;
 * length = best_i + 1; s_last_off = s_off + i where i is the stopping index
 * (n when the walk ran to the end). */
int32_t hbo_extend_right(const int8_t *m, const uint8_t *s, int32_t slen,
                         const uint8_t *q, int32_t qlen, int32_t s_off,
                         int32_t q_off, int32_t dropoff, int32_t maxscore,
                         int32_t *length, int32_t *s_last_off)
{
    int32_t n = slen - s_off < qlen - q_off ? slen - s_off : qlen - q_off;
    int32_t best_i = -1;
    int32_t score = maxscore;
    int32_t i;
    for (i = 0; i < n; ++i) {
        score += m[q[q_off + i] * HBO_ALPHA + s[s_off + i]];
        if (score > maxscore) {
            maxscore = score;
            best_i = i;
        }
        if (score <= 0 || maxscore - score >= dropoff) break;
    }
    *length = best_i + 1;
    *s_last_off = s_off + i;
    return maxscore;
}

/* Two-hit extension (H-BLAST/h_blast_kernel.cu:1069-1130): move the start
 * to just after the best-scoring prefix of the second word, extend left with
 * maxscore 0, and extend right (starting from the left score) only if the
 * left walk reached s_left_off, the end of the first hit's word. */
int32_t hbo_extend_two_hit(const int8_t *m, const uint8_t *s, int32_t slen,
                           const uint8_t *q, int32_t qlen, int32_t s_left_off,
                           int32_t s_right_off, int32_t q_right_off,
                           int32_t dropoff, int32_t *hsp_q, int32_t *hsp_s,
                           int32_t *hsp_len, int32_t *right_extend,
                           int32_t *s_last_off)
{
    int32_t left_d = 0, right_d = 0, left_score = 0, right_score = 0, score = 0;
    for (int32_t i = 0; i < 3; ++i) {
        score += m[q[q_right_off + i] * HBO_ALPHA + s[s_right_off + i]];
        if (score > left_score) {
            left_score = score;
            right_d = i + 1;
        }
    }
    q_right_off += right_d;
    s_right_off += right_d;
    right_d = 0;
    *right_extend = 0;
    *s_last_off = s_right_off;
    left_score = hbo_extend_left(m, s, q, s_right_off - 1, q_right_off - 1,
                                 dropoff, 0, &left_d);
    if (left_d >= s_right_off - s_left_off) {
        *right_extend = 1;
        right_score = hbo_extend_right(m, s, slen, q, qlen, s_right_off,
                                       q_right_off, dropoff, left_score,
                                       &right_d, s_last_off);
    }
    *hsp_q = q_right_off - left_d;
    *hsp_s = s_right_off - left_d;
    *hsp_len = left_d + right_d;
    return left_score > right_score ? left_score : right_score;
}

/* Context of a concatenated-query offset: the last context whose start is
 * <= q_off (binary search as H-BLAST/h_blast_kernel.cu:368-380). */
static int32_t context_of(const hbo_query *q, int32_t q_off)
{
    int32_t lo = 0, hi = q->n_ctx;
    while (hi - lo > 1) {
        int32_t mid = (lo + hi) >> 1;
        if (q->ctx_off[mid] <= q_off) lo = mid;
        else hi = mid;
    }
    return lo;
}

typedef struct hsp_vec {
    hbo_hsp *v;
    int64_t n, cap;
} hsp_vec;

static int vec_push(hsp_vec *h, const hbo_hsp *x)
{
    if (h->n == h->cap) {
        int64_t nc = h->cap ? h->cap * 2 : 64;
        hbo_hsp *nv = (hbo_hsp *)realloc(h->v, (size_t)nc * sizeof(hbo_hsp));
        if (!nv) return HBO_ERR_NOMEM;
        h->v = nv;
        h->cap = nc;
    }
    h->v[h->n++] = *x;
    return 0;
}

typedef struct scan_state {
    int32_t *diag;     /* diag_len entries: bit 31 flag, low 31 bits last_hit + offset */
    int32_t diag_mask;
    int32_t diag_offset;
} scan_state;

/* One subject, sequential in s_off then cell order.  Diagonal state per
 * entry: last_hit + diag_offset and a flag bit set once the diagonal has
 * been extended to the right (H-BLAST/h_blast_kernel.cu:672-870, the
 * diag machine of the CPU two-hit word finder):
 *   flagged:   s < last_hit -> skip; else last_hit = s, flag cleared;
 *   unflagged: diff = s - last_hit; diff >= window -> last_hit = s;
 *              diff < word -> skip; else extend, then
 *              last_hit = right-extended ? s_last_off - (word - 1) (flagged) : s.
 * diag_offset grows by slen + window per subject so stale entries always
 * read as "no recent hit" (H-BLAST/h_blast_kernel.cu:1021). */
static int scan_subject(const hbo_query *q, const hbo_lut *lut, const int8_t *m,
                        const uint8_t *s, int32_t slen, uint32_t oid,
                        scan_state *st, hsp_vec *out, uint32_t *cnt)
{
    const int32_t word = 3, window = q->window, mask = st->diag_mask;
    int32_t *diag = st->diag;
    int32_t off = st->diag_offset;
    uint32_t hits = 0, extended = 0, saved = 0;
    for (int32_t so = 0; so <= slen - word; ++so) {
        const uint32_t idx = ((uint32_t)s[so] << 10) | ((uint32_t)s[so + 1] << 5) | s[so + 2];
        const uint32_t b = lut->start[idx], e = lut->start[idx + 1];
        hits += e - b;
        for (uint32_t k = b; k < e; ++k) {
            const int32_t qo = (int32_t)lut->entries[k];
            const int32_t dc = (qo - so) & mask;
            const int32_t stv = diag[dc];
            const int32_t last = stv & 0x7fffffff;
            if (stv < 0) {
                if (so + off < last) continue;
                diag[dc] = so + off;
                continue;
            }
            const int32_t last_s = last - off;
            const int32_t diff = so - last_s;
            if (diff >= window) {
                diag[dc] = so + off;
                continue;
            }
            if (diff < word) continue;
            const int32_t ctx = context_of(q, qo);
            int32_t hq, hs, hl, rext, slast;
            const int32_t sc = hbo_extend_two_hit(m, s, slen, q->seq, q->len,
                                                  last_s + word, so, qo,
                                                  q->xdrop[ctx], &hq, &hs, &hl,
                                                  &rext, &slast);
            ++extended;
            if (sc >= q->cutoff[ctx]) {
                hbo_hsp h = {oid, ctx, hq, hs, qo, so, hl, sc};
                if (vec_push(out, &h)) return HBO_ERR_NOMEM;
                ++saved;
            }
            if (rext)
                diag[dc] = (int32_t)(0x80000000u | (uint32_t)(slast - (word - 1) + off));
            else
                diag[dc] = so + off;
        }
    }
    if (cnt) {
        cnt[0] = hits;
        cnt[1] = extended;
        cnt[2] = saved;
    }
    /* advance the offset; clear the array before the 31-bit field could wrap */
    if ((int64_t)off + slen + window >= (1 << 30)) {
        memset(diag, 0, (size_t)(mask + 1) * sizeof(int32_t));
        st->diag_offset = window;
    } else {
        st->diag_offset = off + slen + window;
    }
    return 0;
}

typedef struct job {
    const hbo_query *q;
    const hbo_lut *lut;
    const int8_t *m;
    const uint8_t *res;
    const uint64_t *offs;
    uint32_t n;
    uint32_t chunk;
    uint32_t n_chunks;
    uint32_t *counters;
    hsp_vec *chunk_out;    /* one vector per chunk */
    int32_t diag_len;
    volatile uint32_t next;
    volatile int err;
    pthread_mutex_t mu;
} job;

static void *worker(void *arg)
{
    job *J = (job *)arg;
    scan_state st;
    st.diag = (int32_t *)calloc((size_t)J->diag_len, sizeof(int32_t));
    st.diag_mask = J->diag_len - 1;
    st.diag_offset = J->q->window;
    if (!st.diag) {
        J->err = HBO_ERR_NOMEM;
        return NULL;
    }
    for (;;) {
        pthread_mutex_lock(&J->mu);
        uint32_t c = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (c >= J->n_chunks || J->err) break;
        uint32_t lo = c * J->chunk, hi = lo + J->chunk;
        if (hi > J->n) hi = J->n;
        for (uint32_t i = lo; i < hi; ++i) {
            const uint8_t *s = J->res + J->offs[i];
            int32_t slen = (int32_t)(J->offs[i + 1] - J->offs[i]);
            int rc = scan_subject(J->q, J->lut, J->m, s, slen, i, &st,
                                  &J->chunk_out[c],
                                  J->counters ? J->counters + 3 * (size_t)i : NULL);
            if (rc) {
                J->err = rc;
                break;
            }
        }
    }
    free(st.diag);
    return NULL;
}

int64_t hbo_scan_db(const hbo_query *q, const hbo_lut *lut,
                    const uint8_t *residues, const uint64_t *offsets,
                    uint32_t n_subjects, int n_threads, hbo_hsp **out,
                    uint32_t *counters)
{
    int rc = check_query(q);
    if (rc) return rc;
    if (!lut || !lut->start || !residues || !offsets || !out) return HBO_ERR_ARG;
    for (uint32_t i = 0; i < n_subjects; ++i)
        if (offsets[i + 1] < offsets[i]) return HBO_ERR_ARG;
    const uint64_t total = offsets[n_subjects] - offsets[0];
    for (uint64_t i = 0; i < total; ++i)
        if (residues[offsets[0] + i] >= HBO_ALPHA) return HBO_ERR_RESIDUE;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;

    job J;
    memset(&J, 0, sizeof(J));
    J.q = q;
    J.lut = lut;
    J.m = matrix_of(q);
    J.res = residues;
    J.offs = offsets;
    J.n = n_subjects;
    J.chunk = 64;
    J.n_chunks = (n_subjects + J.chunk - 1) / J.chunk;
    J.counters = counters;
    J.diag_len = 1;
    while (J.diag_len < q->len + q->window) J.diag_len <<= 1;
    J.chunk_out = (hsp_vec *)calloc(J.n_chunks ? J.n_chunks : 1, sizeof(hsp_vec));
    if (!J.chunk_out) return HBO_ERR_NOMEM;
    pthread_mutex_init(&J.mu, NULL);

    pthread_t th[256];
    int started = 0;
    if (n_threads == 1) {
        worker(&J);
    } else {
        for (int t = 0; t < n_threads; ++t)
            if (pthread_create(&th[t], NULL, worker, &J) == 0) ++started;
        if (!started) worker(&J);
        for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    }
    pthread_mutex_destroy(&J.mu);

    int64_t n = 0;
    for (uint32_t c = 0; c < J.n_chunks; ++c) n += J.chunk_out[c].n;
    hbo_hsp *all = NULL;
    if (!J.err) {
        all = (hbo_hsp *)malloc((size_t)(n ? n : 1) * sizeof(hbo_hsp));
        if (!all) J.err = HBO_ERR_NOMEM;
    }
    int64_t w = 0;
    for (uint32_t c = 0; c < J.n_chunks; ++c) {
        if (all && J.chunk_out[c].n)
            memcpy(all + w, J.chunk_out[c].v, (size_t)J.chunk_out[c].n * sizeof(hbo_hsp));
        w += J.chunk_out[c].n;
        free(J.chunk_out[c].v);
    }
    free(J.chunk_out);
    if (J.err) {
        free(all);
        return J.err;
    }
    *out = all;
    return n;
}

void hbo_free(void *p) { free(p); }

/* Property check of HSP records against the sequences (for results whose
 * subjects are too many to search with the oracle): each record's subject,
 * context and extent lie inside their sequences, q_off - s_off == q_start -
 * s_start, the score recomputed from the matrix equals the record's and
 * reaches the context's cutoff, and records are in strictly increasing
 * (oid, s_off, q_off) order.  Returns the number of bad records; *first_bad
 * = the first one's index (or n). */
uint64_t hbo_hsp_check(const hbo_query *q, const uint8_t *residues, const uint64_t *offsets,
                       uint32_t n_subjects, const hbo_hsp *h, uint64_t n, uint64_t *first_bad)
{
    const int8_t *m = q->matrix ? q->matrix : hbo_blosum62();
    uint64_t bad = 0;
    *first_bad = n;
    for (uint64_t i = 0; i < n; ++i) {
        const hbo_hsp *r = &h[i];
        int ok = r->oid < n_subjects && r->ctx >= 0 && r->ctx < q->n_ctx && r->len > 0 &&
                 r->q_off - r->s_off == r->q_start - r->s_start;
        if (ok) {
            const int64_t slen = (int64_t)(offsets[r->oid + 1] - offsets[r->oid]);
            const int32_t c0 = q->ctx_off[r->ctx], c1 = c0 + q->ctx_len[r->ctx];
            ok = r->s_start >= 0 && (int64_t)r->s_start + r->len <= slen && r->q_start >= c0 &&
                 r->q_start + r->len <= c1 && r->q_off >= c0 && r->q_off + 3 <= c1 &&
                 r->s_off >= 0 && (int64_t)r->s_off + 3 <= slen;
        }
        if (ok) {
            const uint8_t *s = residues + offsets[r->oid] + r->s_start;
            const uint8_t *qq = q->seq + r->q_start;
            int64_t sc = 0;
            for (int32_t k = 0; k < r->len; ++k) sc += m[qq[k] * HBO_ALPHA + s[k]];
            ok = sc == r->score && r->score >= q->cutoff[r->ctx];
        }
        if (ok && i > 0) {
            const hbo_hsp *p = &h[i - 1];
            ok = p->oid < r->oid || (p->oid == r->oid && (p->s_off < r->s_off ||
                                     (p->s_off == r->s_off && p->q_off < r->q_off)));
        }
        if (!ok) {
            if (!bad) *first_bad = i;
            ++bad;
        }
    }
    return bad;
}
