// hb_gapped.hip -- gfx950 kernels of the preliminary gapped extension.
//
// The consumer of the seed stage: for every subject with ungapped HSPs, the
// in-frame protein path of BLAST_GetGappedScore
// (ncbi_blast_files/blast_gapalign.c:3233-3560), score only:
//   * the subject's HSPs are ranked by (score desc, s_start asc, len desc,
//     q_start asc), the order Blast_InitHitListSortByScore leaves them in
//     (blast_engine.c:584);
//   * an HSP whose start and end points lie in the box of a saved gapped HSP
//     of the same context with at least its score is skipped (the interval
//     tree test, blast_gapalign.c:3402, as a linear scan over the saved list);
//   * otherwise the start point is the middle of its best 11-residue window
//     (BlastGetStartForGappedAlignment, blast_gapalign.c:3184-3229) and the
//     X-drop DP of Blast_SemiGappedAlign (blast_gapalign.c:711-937) runs to
//     the left (through the start pair) and to the right of it
//     (s_BlastProtGappedAlignment, blast_gapalign.c:3654-3783, with
//     AdjustSubjectRange, :3608-3637); the HSP is saved when the two scores
//     sum to the context's cutoff.
//
// Layout on CDNA4: one wavefront owns one subject (subjects are popped 64 at
// a time from a counter).  The DP runs row by row (one query residue per
// row) with the band's columns spread over the 64 lanes, so a row costs a
// few VALU instructions per 64 cells instead of a sequential walk:
//   T(b) = max(diagonal, gap-in-subject score carried down the column)
//   R(b) = best score ending in a gap in the query at b
//        = max over k < b of T(k) - (gap_open + gap_extend) - (b-1-k) gap_extend,
//     an exclusive max-scan of T(k) + k gap_extend;
//   s(b) = max(T(b), R(b)); the cell is dropped when the best score before
//     it (row-major) exceeds s(b) by more than X -- a second max-scan.
// R is computed from every cell, live or dropped.  That equals the
// sequential rule (a dropped cell feeds nothing) because a dropped cell's
// outputs are already below the drop-off threshold, which only rises: any
// cell they would decide is dropped as well, and live cells' values agree.
// The band [lo, end) moves as in the sequential DP (a leading run of
// dropped cells advances lo; a trailing run shortens end; a live last cell
// opens the columns its gap keeps within X; one guard column closes the
// band).  Its columns live in an LDS ring of HB_GAP_RING entries per
// wavefront.  A subject whose band, HSP count or saved list outgrows the
// LDS capacities is handed to a second launch of the same code with its
// ring, order and saved list in global memory.
#include <hip/hip_runtime.h>

#include "hb_internal.h"

namespace {

constexpr int32_t kNeg = INT32_MIN / 2;   // MININT of blast_gapalign.c:61
constexpr int kGapWaves = 4;              // wavefronts per workgroup
constexpr int32_t kWin = 11;              // start-point window (HSP_MAX_WINDOW)
constexpr int32_t kLongSubject = 90000;   // MAX_SUBJECT_OFFSET, blast_gapalign.c:3600
constexpr int32_t kGapAllowance = 3000;   // MAX_TOTAL_GAPS, blast_gapalign.c:3605
static_assert((HB_GAP_RING & (HB_GAP_RING - 1)) == 0, "HB_GAP_RING: power of two");

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int32_t dppn(int32_t v, int32_t old)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dppu(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

// Scores are scanned as v + kBias: every value met (down to kNeg - 128) is
// then positive, so 0 is the identity of an unsigned max and the lane moves
// fuse into v_max_u32 with DPP operands.
constexpr uint32_t kBias = (1u << 30) + (1u << 20);

// Two inclusive max-scans over the wavefront, interleaved (row_shr 1/2/4/8
// within rows of 16 lanes, row_bcast 15/31 across rows).
__device__ __forceinline__ void scan_max2(uint32_t &a, uint32_t &b)
{
    a = max(a, dppu<0x111, 0xf>(a));
    b = max(b, dppu<0x111, 0xf>(b));
    a = max(a, dppu<0x112, 0xf>(a));
    b = max(b, dppu<0x112, 0xf>(b));
    a = max(a, dppu<0x114, 0xf>(a));
    b = max(b, dppu<0x114, 0xf>(b));
    a = max(a, dppu<0x118, 0xf>(a));
    b = max(b, dppu<0x118, 0xf>(b));
    a = max(a, dppu<0x142, 0xa>(a));
    b = max(b, dppu<0x142, 0xa>(b));
    a = max(a, dppu<0x143, 0xc>(a));
    b = max(b, dppu<0x143, 0xc>(b));
}

__device__ __forceinline__ uint32_t scan_max1(uint32_t a)
{
    a = max(a, dppu<0x111, 0xf>(a));
    a = max(a, dppu<0x112, 0xf>(a));
    a = max(a, dppu<0x114, 0xf>(a));
    a = max(a, dppu<0x118, 0xf>(a));
    a = max(a, dppu<0x142, 0xa>(a));
    a = max(a, dppu<0x143, 0xc>(a));
    return a;
}

// Value of lane - 1 (wave_shr:1); lane 0 receives `first`.
__device__ __forceinline__ int32_t from_prev_lane(int32_t v, int32_t first)
{
    return dppn<0x138, 0xf>(v, first);
}

__device__ __forceinline__ uint32_t from_prev_lane_u(uint32_t v)
{
    return dppu<0x138, 0xf>(v);
}

__device__ __forceinline__ int32_t lane_read(int32_t v, int32_t lane)
{
    return __builtin_amdgcn_readlane(v, lane);
}

__device__ __forceinline__ int32_t uniform(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// Lane mask of one compare, and a lane's bit of a mask.  Votes on single
// compares cost nothing beyond the compare, and masks combine in scalar
// instructions; a vote on a combined predicate costs a 0/1 select and a
// compare.
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// Stores of one lane become visible to the other lanes of its wavefront
// (LDS: wavefront scope; the wide variant's global scratch: workgroup scope,
// which waits for the stores to complete).
template <bool WIDE>
__device__ __forceinline__ void sync_lanes()
{
    if (WIDE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

template <bool WIDE>
struct GapWave {
    const hb_gap_args &A;
    const int8_t *mat;     // LDS 32x32
    int2 *ring;            // DP columns: {best score, best ending in a gap in the subject}
    uint32_t rmask;        // ring size - 1
    uint16_t *order16;     // LDS: ranked HSP indices of the current subject
    uint32_t *order32;     // wide pass: the same in global scratch
    int32_t *saved;        // 6 ints per saved HSP: ctx, score, q0, q1, s0, s1
    uint8_t *sbuf;         // LDS copy of the current subject (HB_GAP_SUBJ bytes)
    int32_t lane;
    int32_t go, ge, goe, X;

    __device__ __forceinline__ void put_order(uint32_t r, uint32_t i)
    {
        if (WIDE) order32[r] = i;
        else order16[r] = (uint16_t)i;
    }
    __device__ __forceinline__ uint32_t get_order(uint32_t r) const
    {
        return WIDE ? order32[r] : (uint32_t)order16[r];
    }

    // Ranks the subject's n ungapped HSPs by (score desc, s_start asc, len
    // desc, q_start asc, index asc); order[rank] = index.
    __device__ void rank_hsps(const hb_hsp *U, uint32_t n)
    {
        if (n == 1) {
            if (lane == 0) put_order(0, 0);
            sync_lanes<WIDE>();
            return;
        }
        for (uint32_t i0 = 0; i0 < n; i0 += 64) {
            const uint32_t i = i0 + (uint32_t)lane;
            const bool vi = i < n;
            int32_t sc = 0, ss = 0, ln = 0, qs = 0;
            if (vi) {
                sc = U[i].score;
                ss = U[i].s_start;
                ln = U[i].len;
                qs = U[i].q_start;
            }
            uint32_t rank = 0;
            for (uint32_t j0 = 0; j0 < n; j0 += 64) {
                const uint32_t j = j0 + (uint32_t)lane;
                int32_t ks = 0, kss = 0, kl = 0, kq = 0;
                if (j < n) {
                    ks = U[j].score;
                    kss = U[j].s_start;
                    kl = U[j].len;
                    kq = U[j].q_start;
                }
                const uint32_t nj = n - j0 < 64 ? n - j0 : 64;
                for (uint32_t t = 0; t < nj; ++t) {
                    const int32_t s2 = lane_read(ks, (int32_t)t);
                    const int32_t ss2 = lane_read(kss, (int32_t)t);
                    const int32_t l2 = lane_read(kl, (int32_t)t);
                    const int32_t q2 = lane_read(kq, (int32_t)t);
                    const uint32_t jj = j0 + t;
                    const bool before =
                        s2 > sc ||
                        (s2 == sc &&
                         (ss2 < ss ||
                          (ss2 == ss &&
                           (l2 > ln || (l2 == ln && (q2 < qs || (q2 == qs && jj < i)))))));
                    rank += before ? 1u : 0u;
                }
            }
            if (vi) put_order(rank, i);
        }
        sync_lanes<WIDE>();
    }

    __device__ bool contained(int32_t c, int32_t score, int32_t qs, int32_t ss, int32_t len,
                              uint32_t n_saved) const
    {
        for (uint32_t j0 = 0; j0 < n_saved; j0 += 64) {
            const uint32_t j = j0 + (uint32_t)lane;
            bool hit = false;
            if (j < n_saved) {
                const int32_t *e = saved + 6 * (size_t)j;
                const int32_t q0 = e[2], q1 = e[3], s0 = e[4], s1 = e[5];
                hit = e[0] == c && score <= e[1] && q0 <= qs && q1 >= qs && s0 <= ss &&
                      s1 >= ss && q0 <= qs + len && q1 >= qs + len && s0 <= ss + len &&
                      s1 >= ss + len;
            }
            if (__ballot(hit)) return true;
        }
        return false;
    }

    // Middle of the best 11-residue window of the HSP (first maximum).
    __device__ int32_t start_point(const uint8_t *q, const uint8_t *s, int32_t qs, int32_t len,
                                   int32_t ss) const
    {
        if (len <= kWin) return qs + len / 2;
        int32_t top = kNeg, at = 0;
        for (int32_t e0 = kWin - 1; e0 < len; e0 += 64) {
            const int32_t e = e0 + lane;
            int32_t w = kNeg;
            if (e < len) {
                w = 0;
#pragma unroll
                for (int32_t t = 0; t < kWin; ++t)
                    w += mat[32 * q[qs + e - (kWin - 1) + t] + s[ss + e - (kWin - 1) + t]];
            }
            const int32_t cmax = (int32_t)((uint32_t)lane_read((int32_t)scan_max1((uint32_t)w + kBias), 63) - kBias);
            if (cmax > top) {
                top = cmax;
                at = e0 + (int32_t)__builtin_ctzll(__ballot(e < len && w == cmax));
            }
        }
        return top > 0 ? qs + at - kWin / 2 : qs;
    }

    // Score-only X-drop DP of Blast_SemiGappedAlign from the origin.  Row a
    // (1..M) uses query letter arow[a] (REV: arow[-a]); column b (1..N) uses
    // subject letter subj[so + b] (REV: subj[so + 1 - b]).  Sets *ovf when
    // the band outgrows the ring.
    template <bool REV, class SUBJ>
    __device__ int32_t xdrop(const uint8_t *arow, SUBJ subj, int32_t so, int32_t M, int32_t N,
                             int32_t *a_off, int32_t *b_off, bool *ovf)
    {
        *a_off = 0;
        *b_off = 0;
        if (M <= 0 || N <= 0) return 0;
        const int32_t ring_n = (int32_t)rmask + 1;
        // row 0: leading gaps within the drop-off
        int32_t end = N + 1;
        if (ge > 0) {
            const int32_t k = (X - goe) / ge + 2;
            if (k < end) end = k;
        }
        if (end > ring_n) {
            *ovf = true;
            return 0;
        }
        for (int32_t c0 = 0; c0 < end; c0 += 64) {
            const int32_t c = c0 + lane;
            if (c < end) {
                const int32_t v = c ? -(go + c * ge) : 0;
                ring[c & rmask] = make_int2(v, v - goe);
            }
        }
        sync_lanes<WIDE>();
        const int32_t lane_ge = lane * ge;
        int32_t lo = 0, top = 0, abest = 0, bbest = 0;
        int32_t rows = 0;
        for (int32_t a = 1; a <= M; ++a) {
            if (((a - 1) & 63) == 0) {
                const int32_t ra = a + lane;
                rows = ra <= M ? (int32_t)(REV ? arow[-ra] : arow[ra]) : 0;
            }
            const int8_t *mrow = mat + 32 * lane_read(rows, (a - 1) & 63);
            int32_t last, adv, rafter;
            if (end - lo <= 64) {
                // the whole band in one step (most rows).  Every lane loads
                // and stores (no exec-mask branches): the column is clamped
                // into the subject, and the lanes past the band write
                // (kNeg, kNeg) to ring slots of columns that are rewritten
                // before the band reaches them (the band is under 64
                // columns, so those slots are not the band's)
                const int32_t b = lo + lane;
                const bool valid = b < end;
                const int32_t col = (valid && b >= 1) ? b : 1;
                const int32_t bl = REV ? subj[so + 1 - col] : subj[so + col];
                const int32_t sc = mrow[bl];
                const int2 cell = ring[b & rmask];
                const int32_t ob = valid ? cell.x : kNeg;
                const int32_t ov = valid ? cell.y : kNeg;
                const int32_t pb = from_prev_lane(ob, kNeg);
                const int32_t T = valid ? max(pb + sc, ov) : kNeg;
                uint32_t uA = (uint32_t)(T + lane_ge) + kBias;
                uint32_t uB = (uint32_t)T + kBias;
                scan_max2(uA, uB);
                const uint32_t eA = from_prev_lane_u(uA);
                const uint32_t eB = from_prev_lane_u(uB);
                const int32_t R = (int32_t)(eA - kBias) - goe - lane_ge + ge;
                const int32_t s = max(T, R);
                const int32_t tb = lane ? max(top, (int32_t)(eB - kBias)) : top;
                const uint64_t vm = ballot(valid);
                const uint64_t lm = ballot(tb - s <= X) & vm;
                const bool live = lane_bit(lm);
                ring[b & rmask] = live ? make_int2(s, max(s - goe, ov - ge)) : make_int2(kNeg, ov);
                if (!lm) break;   // every column dropped
                adv = (int32_t)__builtin_ctzll(lm);
                last = lo + 63 - (int32_t)__builtin_clzll(lm);
                const int32_t cmax = (int32_t)((uint32_t)lane_read((int32_t)uB, 63) - kBias);
                if (cmax > top) {
                    top = cmax;
                    abest = a;
                    bbest = lo + (int32_t)__builtin_ctzll(ballot(s == cmax) & vm);
                }
                rafter = lane_read(max(T - goe, R - ge), end - lo - 1);
            } else {
                int32_t carry = kNeg, rin = kNeg, topc = top, rowarg = 0;
                last = -1;
                adv = 0;
                rafter = kNeg;
                bool lead = true;
                for (int32_t c0 = lo; c0 < end; c0 += 64) {
                    const int32_t b = c0 + lane;
                    const bool valid = b < end;
                    const int32_t col = (valid && b >= 1) ? b : 1;
                    const int32_t bl = REV ? subj[so + 1 - col] : subj[so + col];
                    const int32_t sc = mrow[bl];   // loaded by every lane (clamped column)
                    const int2 cell = ring[b & rmask];
                    const int32_t ob = valid ? cell.x : kNeg;
                    const int32_t ov = valid ? cell.y : kNeg;
                    const int32_t pb = from_prev_lane(ob, carry);
                    const int32_t T = valid ? max(pb + sc, ov) : kNeg;
                    uint32_t uA = (uint32_t)(T + lane_ge) + kBias;
                    uint32_t uB = (uint32_t)T + kBias;
                    scan_max2(uA, uB);
                    const uint32_t eA = from_prev_lane_u(uA);
                    const uint32_t eB = from_prev_lane_u(uB);
                    const int32_t P = (int32_t)(eA - kBias);
                    const int32_t R = max(rin - lane_ge, P - goe - lane_ge + ge);
                    const int32_t s = max(T, R);
                    const int32_t before = lane ? max(rin, (int32_t)(eB - kBias)) : kNeg;
                    const int32_t tb = max(topc, before);
                    const uint64_t vm = ballot(valid);
                    const uint64_t lm = ballot(tb - s <= X) & vm;
                    const bool live = lane_bit(lm);
                    if (valid)
                        ring[b & rmask] =
                            live ? make_int2(s, max(s - goe, ov - ge)) : make_int2(kNeg, ov);
                    const int32_t nvalid = end - c0 < 64 ? end - c0 : 64;
                    if (lead) {
                        if (lm) {
                            adv += (int32_t)__builtin_ctzll(lm);
                            lead = false;
                        } else {
                            adv += nvalid;
                        }
                    }
                    if (lm) last = c0 + 63 - (int32_t)__builtin_clzll(lm);
                    const int32_t cmax =
                        max(rin, (int32_t)((uint32_t)lane_read((int32_t)uB, 63) - kBias));
                    if (cmax > topc) {
                        topc = cmax;
                        rowarg = c0 + (int32_t)__builtin_ctzll(ballot(s == cmax) & vm);
                    }
                    const int32_t rnext = max(T - goe, R - ge);   // R after column b
                    rin = lane_read(rnext, 63);
                    rafter = lane_read(rnext, nvalid - 1);
                    carry = lane_read(ob, 63);
                }
                if (topc > top) {
                    top = topc;
                    abest = a;
                    bbest = rowarg;
                }
                if (lo + adv == end) break;   // every column dropped
            }
            lo += adv;
            int32_t new_end, n_ext = 0;
            if (last < end - 1) {
                new_end = last + 1;
            } else {
                const int32_t thr = top - X;
                if (rafter >= thr) {
                    n_ext = N + 1 - end;
                    if (ge > 0) {
                        // no division for the usual gap extension of 1
                        const uint32_t d = (uint32_t)(rafter - thr);
                        const int32_t k = (int32_t)(ge == 1 ? d : d / (uint32_t)ge) + 1;
                        if (k < n_ext) n_ext = k;
                    }
                }
                new_end = end + n_ext;
            }
            const int32_t guard = new_end <= N ? 1 : 0;
            if (new_end + guard - lo > ring_n) {
                *ovf = true;
                return 0;
            }
            for (int32_t i0 = 0; i0 < n_ext; i0 += 64) {
                const int32_t i = i0 + lane;
                if (i < n_ext) {
                    const int32_t v = rafter - i * ge;
                    ring[(end + i) & rmask] = make_int2(v, v - goe);
                }
            }
            // every lane stores the same guard cell (a uniform branch, no
            // single-lane exec mask)
            if (guard) ring[new_end & rmask] = make_int2(kNeg, kNeg);
            end = new_end + guard;
            sync_lanes<WIDE>();
        }
        *a_off = abest;
        *b_off = bbest;
        return top;
    }

    // Both sides of one extension through (qo, so); returns the score and
    // fills the ranges.
    template <class SUBJ>
    __device__ int32_t extend(const uint8_t *qc, int32_t clen, SUBJ subj, int32_t slen,
                              int32_t qo, int32_t so, int32_t *co, bool *ovf)
    {
        // AdjustSubjectRange
        int32_t ql = qo + 1, sl = so + 1, sub_len = slen, shift = 0;
        if (slen >= kLongSubject) {
            const int32_t reach_left = ql + kGapAllowance;
            const int32_t reach_right = clen - ql + kGapAllowance;
            const int32_t s_off1 = sl;
            if (s_off1 > reach_left) {
                shift = s_off1 - reach_left;
                sl = reach_left;
            }
            sub_len = min(slen, s_off1 + reach_right) - shift;
        }
        int32_t a, b;
        const int32_t left = xdrop<true>(qc + qo + 1, subj, so, ql, sl, &a, &b, ovf);
        co[0] = ql - a;
        co[2] = sl - b + shift;
        int32_t right = 0;
        co[1] = qo;
        co[3] = so;
        if (ql < clen && sl < sub_len) {
            right = xdrop<false>(qc + qo, subj, so, clen - ql, sub_len - sl, &a, &b, ovf);
            co[1] = a + qo + 1;
            co[3] = b + so + 1;
        }
        return left + right;
    }

    // One subject; false when it outgrows the LDS capacities.
    __device__ bool subject(uint32_t k)
    {
        const hb_subject_stats st = A.ust[k];
        const uint32_t n = st.n_hsp;
        if (!WIDE && n > HB_GAP_ORDER) return false;
        const uint64_t f = (uint64_t)st.first;
        const uint64_t sbase = A.offs[k];
        const int32_t slen = (int32_t)(A.offs[k + 1] - sbase);
        const uint8_t *subj = A.res + sbase;
        const hb_hsp *U = A.urec + f;
        const bool staged = !WIDE && slen <= HB_GAP_SUBJ;
        if (staged) {
            // whole aligned words around the subject, 4 bytes per lane per
            // step (the residue buffer has 64 bytes of padding; bytes outside
            // the subject are never used)
            const uint32_t *src = reinterpret_cast<const uint32_t *>(subj - (sbase & 3));
            uint32_t *dst = reinterpret_cast<uint32_t *>(sbuf);
            const int32_t words = (int32_t)((slen + (sbase & 3) + 3) >> 2);
            for (int32_t w = lane; w < words; w += 64) dst[w] = src[w];
        }
        const int32_t sofs = staged ? (int32_t)(sbase & 3) : 0;
        rank_hsps(U, n);
        uint32_t n_saved = 0, ext = 0;
        for (uint32_t r = 0; r < n; ++r) {
            const uint32_t i = (uint32_t)uniform((int32_t)get_order(r));
            const int32_t c = uniform(U[i].context);
            const int32_t base = A.ctx[c];
            const int32_t clen = A.ctx[A.n_ctx + c];
            const int32_t cut = A.ctx[2 * A.n_ctx + c];
            const uint8_t *qc = A.query + base;
            const int32_t score0 = uniform(U[i].score);
            const int32_t qs = uniform(U[i].q_start) - base;
            const int32_t ss = uniform(U[i].s_start);
            const int32_t len = uniform(U[i].len);
            if (contained(c, score0, qs, ss, len, n_saved)) continue;
            ++ext;
            const int32_t qo = start_point(qc, subj, qs, len, ss);
            const int32_t so = uniform(U[i].s_off) + (qo - (uniform(U[i].q_off) - base));
            bool ovf = false;
            int32_t co[4];
            int32_t score;
            if (staged)
                score = extend(qc, clen, (const uint8_t *)sbuf + sofs, slen, qo, so, co, &ovf);
            else
                score = extend(qc, clen, subj, slen, qo, so, co, &ovf);
            if (ovf) return false;
            if (score >= cut) {
                if (!WIDE && n_saved >= HB_GAP_SAVED) return false;
                if (lane == 0) {
                    hb_gapped_hsp g;
                    g.oid = k;
                    g.context = c;
                    g.q_start = co[0];
                    g.q_end = co[1];
                    g.s_start = co[2];
                    g.s_end = co[3];
                    g.q_gapped_start = qo;
                    g.s_gapped_start = so;
                    g.score = score;
                    g.pad = 0;
                    A.out[f + n_saved] = g;
                    int32_t *e = saved + 6 * (size_t)n_saved;
                    e[0] = c;
                    e[1] = score;
                    e[2] = co[0];
                    e[3] = co[1];
                    e[4] = co[2];
                    e[5] = co[3];
                }
                sync_lanes<WIDE>();
                ++n_saved;
            }
        }
        if (lane == 0) {
            A.st3[3 * (size_t)k] = 0;
            A.st3[3 * (size_t)k + 1] = ext;
            A.st3[3 * (size_t)k + 2] = n_saved;
        }
        return true;
    }
};

// Per-wavefront LDS of the main pass.
struct GapLds {
    int2 ring[HB_GAP_RING];
    int32_t saved[6 * HB_GAP_SAVED];
    uint16_t order[HB_GAP_ORDER];
    uint32_t subj[HB_GAP_SUBJ / 4 + 2];   // whole words around the subject
};

template <bool WIDE>
__global__ __launch_bounds__(kGapWaves * HB_WAVE) void gapped_kernel(hb_gap_args A)
{
    __shared__ int8_t mat[32 * 32];
    __shared__ GapLds lds[WIDE ? 1 : kGapWaves];
    for (int i = threadIdx.x; i < 32 * 32 / 4; i += blockDim.x)
        reinterpret_cast<int32_t *>(mat)[i] = reinterpret_cast<const int32_t *>(A.mat32)[i];
    __syncthreads();
    const int wave = threadIdx.x / HB_WAVE;
    const int lane = threadIdx.x % HB_WAVE;
    GapWave<WIDE> W{A, mat, nullptr, 0, nullptr, nullptr, nullptr, nullptr, lane, 0, 0, 0, 0};
    W.go = A.gap_open;
    W.ge = A.gap_extend;
    W.goe = A.gap_open + A.gap_extend;
    W.X = A.x_dropoff < W.goe ? W.goe : A.x_dropoff;
    if (!WIDE) {
        GapLds &L = lds[wave];
        W.ring = L.ring;
        W.rmask = HB_GAP_RING - 1;
        W.order16 = L.order;
        W.saved = L.saved;
        W.sbuf = reinterpret_cast<uint8_t *>(L.subj);
        // one subject per pop, subjects with the most HSPs first
        const uint32_t n_list = (uint32_t)uniform((int32_t)*A.n_list);
        for (;;) {
            uint32_t item = 0;
            if (lane == 0) item = atomicAdd(A.counter, 1u);
            item = (uint32_t)lane_read((int32_t)item, 0);
            if (item >= n_list) break;
            const uint32_t k = (uint32_t)uniform((int32_t)A.list[item]);
            if (!W.subject(k) && lane == 0) {
                const uint32_t p = atomicAdd(A.wide, 1u);
                A.wide[1 + p] = k;
            }
        }
    } else {
        const size_t gw = (size_t)blockIdx.x * kGapWaves + wave;
        int32_t *scr = A.scratch + gw * A.scratch_words;
        W.ring = reinterpret_cast<int2 *>(scr);
        W.rmask = A.ring_wide - 1;
        W.saved = scr + 2 * (size_t)A.ring_wide;
        W.order32 = reinterpret_cast<uint32_t *>(W.saved + 6 * (size_t)A.n_max);
        for (;;) {
            uint32_t item = 0;
            if (lane == 0) item = atomicAdd(A.counter, 1u);
            item = (uint32_t)lane_read((int32_t)item, 0);
            if (item >= A.n_wide) break;
            W.subject(A.wide[1 + item]);
        }
    }
}

// Work list: bucket b = 63 - min(n_hsp, 63) for subjects with HSPs, so that
// ascending buckets are descending HSP counts.
__device__ __forceinline__ uint32_t gap_bucket_of(uint32_t n) { return 63 - (n < 63 ? n : 63); }

__global__ __launch_bounds__(256) void gap_count_kernel(hb_gap_args A, uint32_t *bucket)
{
    __shared__ uint32_t h[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < A.n_subjects) {
        const uint32_t n = A.ust[k].n_hsp;
        if (n == 0) {
            A.st3[3 * (size_t)k] = 0;
            A.st3[3 * (size_t)k + 1] = 0;
            A.st3[3 * (size_t)k + 2] = 0;
        } else {
            atomicAdd(&h[gap_bucket_of(n)], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x]) atomicAdd(&bucket[threadIdx.x], h[threadIdx.x]);
}

// One wavefront: exclusive prefix of the 64 bucket counts into the cursors
// (bucket[64 ..]) and the total.
__global__ void gap_bucket_kernel(uint32_t *bucket, uint32_t *n_list)
{
    const uint32_t lane = threadIdx.x;
    const uint32_t v = bucket[lane];
    uint32_t incl = v;
    for (uint32_t d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    bucket[64 + lane] = incl - v;
    if (lane == 63) *n_list = incl;
}

// Places each subject with HSPs in its bucket: a workgroup reserves one range
// per bucket with a single global atomic, its threads take slots in LDS.
__global__ __launch_bounds__(256) void gap_place_kernel(hb_gap_args A, uint32_t *bucket,
                                                        uint32_t *list)
{
    __shared__ uint32_t h[64], base[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0;
    __syncthreads();
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t bk = 64, slot = 0;
    if (k < A.n_subjects) {
        const uint32_t n = A.ust[k].n_hsp;
        if (n) {
            bk = gap_bucket_of(n);
            slot = atomicAdd(&h[bk], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x])
        base[threadIdx.x] = atomicAdd(&bucket[64 + threadIdx.x], h[threadIdx.x]);
    __syncthreads();
    if (bk < 64) list[base[bk] + slot] = k;
}

__global__ void count_oids_kernel(const hb_hsp *rec, uint32_t n, uint32_t *st3)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&st3[3 * (size_t)rec[i].oid + 2], 1u);
}

// One thread per subject: dense layout of the gapped HSPs and the stats.
__global__ void gap_compact_kernel(hb_gap_args A, const uint32_t *first, hb_gapped_stats *stats,
                                   hb_gapped_hsp *dense)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= A.n_subjects) return;
    const uint32_t n = A.st3[3 * (size_t)k + 2];
    hb_gapped_stats o;
    o.extensions = A.st3[3 * (size_t)k + 1];
    o.n_hsp = n;
    o.first = n ? (int64_t)first[k] : -1;
    stats[k] = o;
    if (n) {
        const hb_gapped_hsp *src = A.out + A.ust[k].first;
        for (uint32_t i = 0; i < n; ++i) dense[first[k] + i] = src[i];
    }
}

}  // namespace

hipError_t hb_launch_gap_list(const hb_gap_args &a, uint32_t *list, uint32_t *n_list,
                              uint32_t *scratch, hipStream_t st)
{
    hipError_t e = hb_launch_fill_u32(scratch, 0u, 128, st);
    if (e != hipSuccess || a.n_subjects == 0) return e;
    const dim3 grid((a.n_subjects + 255) / 256);
    hipLaunchKernelGGL(gap_count_kernel, grid, dim3(256), 0, st, a, scratch);
    hipLaunchKernelGGL(gap_bucket_kernel, dim3(1), dim3(64), 0, st, scratch, n_list);
    hipLaunchKernelGGL(gap_place_kernel, grid, dim3(256), 0, st, a, scratch, list);
    return hipGetLastError();
}

hipError_t hb_launch_count_oids(const hb_hsp *rec, uint32_t n, uint32_t *st3, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(count_oids_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rec, n, st3);
    return hipGetLastError();
}

hipError_t hb_launch_gapped(const hb_gap_args &a, bool wide, int n_blocks, hipStream_t st)
{
    if (n_blocks < 1) return hipSuccess;
    if (wide)
        hipLaunchKernelGGL(gapped_kernel<true>, dim3(n_blocks), dim3(kGapWaves * HB_WAVE), 0, st,
                           a);
    else
        hipLaunchKernelGGL(gapped_kernel<false>, dim3(n_blocks), dim3(kGapWaves * HB_WAVE), 0,
                           st, a);
    return hipGetLastError();
}

hipError_t hb_launch_gap_compact(const hb_gap_args &a, const uint32_t *first,
                                 hb_gapped_stats *stats, hb_gapped_hsp *out_dense,
                                 hipStream_t st)
{
    if (a.n_subjects == 0) return hipSuccess;
    hipLaunchKernelGGL(gap_compact_kernel, dim3((a.n_subjects + 255) / 256), dim3(256), 0, st, a,
                       first, stats, out_dense);
    return hipGetLastError();
}
