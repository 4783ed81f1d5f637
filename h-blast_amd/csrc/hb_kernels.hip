// hb_kernels.hip -- gfx950 kernels of the seed-extension stage.
//
// Scan kernel (hb_scan_kernel): the BLASTP/BLASTX word scan, two-hit
// diagonal tracking and X-drop ungapped extension of
// h_blast_v1.1/H-BLAST/h_blast_kernel.cu:21-1540, restructured for CDNA4:
//   * one wavefront owns one subject at a time (staged into LDS when it
//     fits); subjects come from a length-descending tile queue;
//   * lane i owns subject position s0+i of a 64-position window and fetches
//     the cell {start, count} of its 3-mer from the lookup table (L2
//     resident); a wavefront prefix sum flattens the window's hits in
//     (s, q) order;
//   * hits are decided 64 per round in one pass: the per-diagonal state
//     {last_hit + offset | flag | pending} lives in one 32-bit LDS word per
//     diagonal whose top bits order the hits of a round that share the
//     diagonal (an atomic max of 64 - lane; the lowest lane decides), and a
//     hit that loses is carried to the next round ahead of newer hits, so
//     every diagonal sees its hits in subject order -- the exact sequential
//     semantics of the CPU two-hit word finder (the reference kernel's
//     deferral, h_blast_kernel.cu:873-942, is not reproduced);
//   * extensions are queued and run 64 at a time, one per lane, as
//     branch-free X-drop walks over LDS data; subjects stay staged in a
//     per-wavefront ring until their extensions ran, so a batch collects
//     the extensions of several subjects; hits on diagonals whose extension
//     is pending wait in an ordered list and are replayed after.
#include <hip/hip_runtime.h>

#include "hb_internal.h"

// Global (address space 1) pointee qualifier.
#define HB_GLOBAL __attribute__((address_space(1)))

namespace {

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-wide inclusive scans on DPP lane moves (VALU only): row_shr
// 1/2/4/8 scans each row of 16 lanes, row_bcast:15 and row_bcast:31 carry the
// row totals into the following rows.  Lanes outside a row read 0, the
// identity of both operations on unsigned values.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ uint32_t dpp(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROW_MASK, 0xf, false);
}

__device__ __forceinline__ uint32_t wave_incl_add(uint32_t v)
{
    v += dpp<0x111, 0xf>(v);  // row_shr:1
    v += dpp<0x112, 0xf>(v);  // row_shr:2
    v += dpp<0x114, 0xf>(v);  // row_shr:4
    v += dpp<0x118, 0xf>(v);  // row_shr:8
    v += dpp<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return v;
}

__device__ __forceinline__ uint32_t wave_incl_max(uint32_t v)
{
    v = max(v, dpp<0x111, 0xf>(v));
    v = max(v, dpp<0x112, 0xf>(v));
    v = max(v, dpp<0x114, 0xf>(v));
    v = max(v, dpp<0x118, 0xf>(v));
    v = max(v, dpp<0x142, 0xa>(v));
    v = max(v, dpp<0x143, 0xc>(v));
    return v;
}

__device__ __forceinline__ uint32_t lane63(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// Lane mask of a compare, and a lane's bit of a mask (no instructions when
// the mask is used as the lane predicate).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool lane_bit(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// ---------------------------------------------------------------- residues
__global__ void check_residues_kernel(const uint8_t *res, uint64_t n, uint32_t *flag)
{
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t bad = 0;
    const uint64_t n4 = n / 4;
    const uint32_t *r4 = reinterpret_cast<const uint32_t *>(res);
    for (uint64_t k = i; k < n4; k += stride) {
        const uint32_t w = r4[k];
#pragma unroll
        for (int b = 0; b < 4; ++b) bad |= (((w >> (8 * b)) & 0xffu) >= HB_ALPHA);
    }
    for (uint64_t k = n4 * 4 + i; k < n; k += stride) bad |= (res[k] >= HB_ALPHA);
    if (bad) atomicOr(flag, 1u);
}

// ---------------------------------------------------------------- lookup table
// Cell w = (a<<10)|(b<<5)|c holds query offset q (a word inside one context)
// when M[q0][a]+M[q1][b]+M[q2][c] >= T, or when w is the query word itself
// (an exact word always seeds, also below T).  One thread per cell, query
// offsets visited ascending, so each cell lists them ascending (as
// lut_sorting leaves them, h_blast.cu:165-194).  grid.y = group.
template <int PASS>
__global__ __launch_bounds__(256) void lut_kernel(const hb_lut_job *jobs, const int8_t *mat32,
                                                  int32_t threshold)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const hb_lut_job J = jobs[blockIdx.y];
    int8_t *sm = reinterpret_cast<int8_t *>(lds);        // 32x32
    uint8_t *sq = lds + 1024;                              // qlen bytes
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) sm[i] = mat32[i];
    for (int i = threadIdx.x; i < J.qlen; i += blockDim.x) sq[i] = J.qflag[i];
    __syncthreads();
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t a = w >> 10, b = (w >> 5) & 31, c = w & 31;
    uint32_t n = 0;
    if (a < HB_ALPHA && b < HB_ALPHA && c < HB_ALPHA) {
        uint32_t k = PASS == 2 ? J.count_start[w] : 0;
        for (int32_t q = 0; q + 3 <= J.qlen; ++q) {
            const uint8_t f0 = sq[q];
            if (!(f0 & 0x80)) continue;
            const uint32_t q0 = f0 & 31, q1 = sq[q + 1] & 31, q2 = sq[q + 2] & 31;
            const int32_t sc = sm[(q0 << 5) | a] + sm[(q1 << 5) | b] + sm[(q2 << 5) | c];
            const bool self = (q0 == a) & (q1 == b) & (q2 == c);
            if (sc >= threshold || self) {
                if (PASS == 2) J.entries[k++] = (uint32_t)q;
                ++n;
            }
        }
    }
    if (PASS == 0) {
        J.count_start[w] = n;
    } else {
        J.cells[w] = make_uint2(J.count_start[w], n);
        const uint64_t m = __ballot(n != 0);
        const int lane = threadIdx.x & 63;
        if (lane == 0) J.pv[w >> 5] = (uint32_t)m;
        if (lane == 32) J.pv[w >> 5] = (uint32_t)(m >> 32);
    }
}

// Exclusive scan of one group's HB_NCELLS counts in place; [HB_NCELLS] and
// *total receive the number of entries.  grid.y = group.
__global__ __launch_bounds__(1024) void lut_scan_kernel(const hb_lut_job *jobs)
{
    __shared__ uint32_t part[1024];
    const hb_lut_job J = jobs[blockIdx.y];
    uint32_t *cnt = J.count_start;
    const int t = threadIdx.x;
    const int per = HB_NCELLS / 1024;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i) s += cnt[t * per + i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const uint32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = 0; i < per; ++i) {
        const uint32_t v = cnt[t * per + i];
        cnt[t * per + i] = run;
        run += v;
    }
    if (t == 1023) {
        cnt[HB_NCELLS] = run;
        *J.total = run;
    }
}

// ---------------------------------------------------------------- extensions
// Left X-drop walk from (s_off, q_off) down to offset 0
// (h_blast_kernel.cu:1329-1540): positions k = 0..n, n = min(s_off, q_off);
// strict '>' keeps the first best; stops once max - score >= X.  Returns the
// best score; *len = positions kept (0 if no prefix improved on maxscore).
// The walk is branch-free per lane: eight residue pairs are fetched at once,
// a lane that stopped keeps its best score and length (its running score
// may change afterwards and is not used), and the loop ends when no lane of
// the wavefront is still walking.  With the subject in LDS (SL) the fetches
// past a sequence end read other LDS bytes (or 0 past the allocation) and
// are ignored; with the subject in global memory the positions are clamped
// into the sequences.
template <bool SL, typename SP>
__device__ __forceinline__ int32_t ext_left(const int8_t *m, SP S, const uint8_t *Q,
                                            int32_t s_off, int32_t q_off, int32_t X,
                                            int32_t maxscore, int32_t *len)
{
    const int32_t n = min(s_off, q_off);
    int32_t best = -1, sc = maxscore;
    bool done = n < 0;
    for (int32_t k0 = 0;; k0 += 8) {
        int32_t mv[8];
        if (SL) {
            // qp[7 - j] = Q[q_off - k0 - j]
            const uint8_t *qp = Q + (q_off - k0 - 7);
            const SP sp = S + (s_off - k0 - 7);
#pragma unroll
            for (int j = 0; j < 8; ++j) mv[j] = m[((uint32_t)qp[7 - j] << 5) | sp[7 - j]];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int32_t k = max(min(k0 + j, n), 0);
                mv[j] = m[((uint32_t)Q[q_off - k] << 5) | S[s_off - k]];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool live = !done && k0 + j <= n;
            sc += mv[j];
            const bool imp = live && sc > maxscore;
            maxscore = imp ? sc : maxscore;
            best = imp ? k0 + j : best;
            done = !live || maxscore - sc >= X;
        }
        if (!__ballot(!done)) break;
    }
    *len = best + 1;
    return maxscore;
}

// Right X-drop walk (h_blast_kernel.cu:1133-1326): positions i = 0..n-1,
// n = min(slen - s_off, qlen - q_off); also stops at score <= 0;
// *s_last = s_off + stopping index (n if the walk ran out).  Branch-free
// per lane like ext_left.
template <bool SL, typename SP>
__device__ __forceinline__ int32_t ext_right(const int8_t *m, SP S, int32_t slen, const uint8_t *Q,
                                             int32_t qlen, int32_t s_off, int32_t q_off, int32_t X,
                                             int32_t maxscore, int32_t *len, int32_t *s_last)
{
    const int32_t n = min(slen - s_off, qlen - q_off);
    int32_t best = -1, sc = maxscore, stop = n;
    bool done = n <= 0;
    for (int32_t i0 = 0;; i0 += 8) {
        int32_t mv[8];
        if (SL) {
            const uint8_t *qp = Q + (q_off + i0);
            const SP sp = S + (s_off + i0);
#pragma unroll
            for (int j = 0; j < 8; ++j) mv[j] = m[((uint32_t)qp[j] << 5) | sp[j]];
        } else {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int32_t i = max(min(i0 + j, n - 1), 0);
                mv[j] = m[((uint32_t)Q[q_off + i] << 5) | S[s_off + i]];
            }
        }
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const bool live = !done && i0 + j < n;
            sc += mv[j];
            const bool imp = live && sc > maxscore;
            maxscore = imp ? sc : maxscore;
            best = imp ? i0 + j : best;
            const bool brk = live && (sc <= 0 || maxscore - sc >= X);
            stop = brk ? i0 + j : stop;
            done = !live || brk;
        }
        if (!__ballot(!done)) break;
    }
    *len = best + 1;
    *s_last = s_off + stop;
    return maxscore;
}

struct ext_result {
    int32_t q_start, s_start, len, score, right, s_last;
};

// Two-hit extension (h_blast_kernel.cu:1069-1130): start just after the
// best-scoring prefix of the second word, extend left from maxscore 0, and
// extend right from the left score only if the left walk reached
// s_left_off (the end of the first hit's word).
template <bool SL, typename SP>
__device__ __forceinline__ ext_result ext_two_hit(const int8_t *m, SP S, int32_t slen,
                                                  const uint8_t *Q, int32_t qlen,
                                                  int32_t s_left_off, int32_t s_right_off,
                                                  int32_t q_right_off, int32_t X)
{
    int32_t sc = 0, left_score = 0, right_d = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        sc += m[((uint32_t)Q[q_right_off + i] << 5) | S[s_right_off + i]];
        if (sc > left_score) {
            left_score = sc;
            right_d = i + 1;
        }
    }
    q_right_off += right_d;
    s_right_off += right_d;
    ext_result r;
    int32_t left_len = 0, right_len = 0, right_score = 0;
    r.right = 0;
    r.s_last = s_right_off;
    left_score = ext_left<SL>(m, S, Q, s_right_off - 1, q_right_off - 1, X, 0, &left_len);
    if (left_len >= s_right_off - s_left_off) {
        r.right = 1;
        right_score = ext_right<SL>(m, S, slen, Q, qlen, s_right_off, q_right_off, X, left_score,
                                &right_len, &r.s_last);
    }
    r.q_start = q_right_off - left_len;
    r.s_start = s_right_off - left_len;
    r.len = left_len + right_len;
    r.score = max(left_score, right_score);
    return r;
}

__device__ __forceinline__ int32_t context_of(const int32_t *ctx_off, int32_t n_ctx, int32_t q)
{
    int32_t lo = 0, hi = n_ctx;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (ctx_off[mid] <= q) lo = mid;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------- scan
// Per-wavefront LDS of the scan (D = diagonal array length of the launch):
//   word[D]    uint32 per diagonal: bits 0..22 last hit + offset, bit 23
//              extension pending, bit 24 extended-right flag, bits 25..31
//              the ordering marker of the decision pass under way (0
//              between passes).  With the state in global memory (long
//              queries) the word is a uint32 in the wavefront's own region
//              (29-bit position, bit 29 pending, bit 30 flag) and the
//              markers sit in a table of HB_GMARKS words here instead,
//              hashed by diagonal.
//   seg[64]    uint32 tagged segment starts of the current hit round
//   jobs[kJobCap] uint2 queued two-hit extensions (a ring, oldest first):
//              (q | (s - last) << 24, s | slot << 28)
//   blk[kBlkCap]  uint2 hits waiting on a pending diagonal (q, s | slot << 28)
//   slot[HB_SLOTS] per staged subject whose extensions may still be queued
//              (a ring, oldest first): {oid, length, diagonal offset, ring
//              offset} and {lookup hits, extensions, HSPs, 0}
//   ring[ring] residues of those subjects (16-byte aligned, circular: a
//              subject that does not fit before the end starts at 0); the
//              oldest subject leaves once none of its extensions is queued
// Per-workgroup LDS: score matrix, context tables (stride ctx_pad_max) and
// query of the group being scanned.
constexpr uint32_t kJobCap = HB_JOBCAP;
constexpr uint32_t kGMarks = HB_GMARKS;
static_assert((kGMarks & (kGMarks - 1)) == 0, "marker table");
constexpr uint32_t kBlkCap = HB_BLKCAP;
static_assert((kJobCap & (kJobCap - 1)) == 0 && kJobCap >= 2 * HB_WAVE, "job ring");
static_assert(kBlkCap >= 2 * HB_WAVE, "blocked list");
static_assert(HB_SLOTS <= 8, "slot field: 3 bits");
constexpr uint32_t kSoMask = 0x0fffffffu;   // subject offset in job and blocked records

__device__ __forceinline__ uint32_t lane_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Diagonal word formats: marker shift, flag, pending, position mask.  The
// state is a position p < kP (idle, last hit p), kP (an extension pending)
// or kF | p (extended to the right up to p); kF = 2 kP and 3 kP < 2^31, so
// for a hit at offset sb < kP the signed difference sb - state is sb - p
// when idle, in [-kP, 0) when pending and sb - p - kF when flagged.
template <bool G> struct DiagWord;
template <> struct DiagWord<false> {
    using T = uint32_t;
    static constexpr int kMS = 25;
    static constexpr uint32_t kF = 1u << 24, kP = 1u << 23, kSM = (1u << 25) - 1;
};
template <> struct DiagWord<true> {   // no marker bits: the markers are in an LDS table
    using T = uint32_t;
    static constexpr uint32_t kF = 1u << 30, kP = 1u << 29, kSM = 0xffffffffu;
};
static_assert(HB_LDS_POS_LIMIT <= (1 << 23) && HB_GLOBAL_POS_LIMIT <= (1 << 28),
              "positions (< limit + longest subject) must stay below kP");

template <bool G>
struct WaveScan {
    using W = DiagWord<G>;
    using word_t = typename W::T;
    const hb_scan_args *a;
    // group being scanned
    // Lookup-table pointers come from a descriptor in memory, so the
    // compiler cannot tell their address space; as generic pointers their
    // loads would be flat loads, which also count against the LDS wait
    // counter.
    const HB_GLOBAL uint64_t *cells;   // {start, count} as one 64-bit word
    const HB_GLOBAL uint32_t *entries;
    uint32_t n_entries;
    int32_t qlen;
    int32_t n_ctx;
    int32_t ctx_base;
    int32_t D;
    uint32_t dmul;     // ceil(2^32 / D): x mod D by a multiply-high
    // LDS tables
    const int8_t *mat;
    const uint8_t *qry;
    const int32_t *ctx;   // ctx_off[ctx_pad], xdrop[ctx_pad], cutoff[ctx_pad]
    int32_t ctx_pad;
    word_t *wd;        // diagonal words (LDS, or the wavefront's global region)
    uint32_t *mkt;     // global-memory words: the wavefront's LDS marker table
    uint32_t *seg;
    uint2 *jobs;
    uint2 *blk;
    uint4 *sla;        // per slot {oid, length, diagonal offset, ring offset}
    uint4 *slb;        // per slot {lookup hits, extensions, HSPs, 0}
    uint8_t *ring;
    int lane;
    int32_t window;
    int32_t base;      // diagonal offset of the next subject: stale entries read as "no recent hit"
    uint32_t rgen;     // segment-tag generation
    uint32_t j_head;   // oldest queued extension (wave-uniform)
    uint32_t n_jobs;   // wave-uniform
    uint32_t n_blk;    // wave-uniform
    uint32_t n_live;   // staged subjects not yet retired (slots s_tail ..)
    uint32_t s_tail;   // slot of the oldest staged subject
    uint32_t ring_head;  // ring offset after the newest staged subject
    bool rwrap;        // the staged residues wrap: newer subjects start at 0
    // running totals, per lane (a lane adds the counters of the subjects it
    // retires), so that they take no scalar registers in the scan loops
    uint64_t tot_hits, tot_ext, tot_hsp;

    __device__ __forceinline__ uint32_t mod_d(uint32_t x) const
    {
        // x < 2^28, D < 2^25: the estimate is the quotient or one more
        const uint32_t qe = __umulhi(x, dmul);
        const int32_t r = (int32_t)(x - qe * (uint32_t)D);
        return (uint32_t)(r < 0 ? r + D : r);
    }
    __device__ __forceinline__ uint32_t diag_of(uint32_t q, int32_t so) const
    {
        const int32_t d = (int32_t)q - (int32_t)mod_d((uint32_t)so);
        return (uint32_t)(d < 0 ? d + D : d);
    }

    // Diagonal word accesses.  G = false: LDS.  G = true: the wavefront's
    // own global-memory region; the accesses are workgroup-scope atomics and
    // sync() waits for them, so a lane reads what another lane of the wave
    // wrote before the last sync().
    __device__ __forceinline__ word_t w_max(uint32_t d, word_t v)
    {
        if constexpr (G)
            return __hip_atomic_fetch_max(&wd[d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            return atomicMax(&wd[d], v);
    }
    __device__ __forceinline__ void w_or(uint32_t d, word_t v)
    {
        if constexpr (G)
            __hip_atomic_fetch_or(&wd[d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            atomicOr(&wd[d], v);
    }
    __device__ __forceinline__ word_t w_ld(uint32_t d) const
    {
        if constexpr (G)
            return __hip_atomic_load(&wd[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            return wd[d];
    }
    __device__ __forceinline__ void w_st(uint32_t d, word_t v)
    {
        if constexpr (G)
            __hip_atomic_store(&wd[d], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
            wd[d] = v;
    }
    __device__ __forceinline__ void sync() const
    {
        if constexpr (G) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
        wave_lds_sync();
    }

    // Clears the diagonal words (a new group, or positions near the limit).
    __device__ __forceinline__ void clear_words()
    {
        for (int i = lane; i < D; i += HB_WAVE) w_st((uint32_t)i, 0);
        sync();
        base = window;
    }

    // One pass of two-hit decisions over a set of hits (lane mask am, at
    // most one hit per lane; the set's lanes are in subject order).  Among
    // the hits of the set that share a diagonal only the lowest lane
    // decides: each hit raises the marker bits of its diagonal's word to
    // 64 - lane with an atomic max (which returns the state to the first hit
    // to arrive and clears the state bits), the hit that finds its own marker
    // after the pass's sync decides, and its store of the new state clears
    // the marker.  The winner is usually the first to arrive; when it is
    // not, the first hit ORs the state back into the word for it.  A winner
    // whose diagonal waits on a pending extension joins the blocked list.
    // The losers then read the winner's new state: a loser whose diagonal is
    // now pending joins the blocked list behind the winner, and a loser that
    // would skip against that state is done (a skip changes nothing, and
    // skipping is monotone along a diagonal: a hit skips a flagged state iff
    // s < last, an unflagged one iff s - last < 3, so every earlier loser of
    // the diagonal skips too).  Returns the mask of the other losers: they
    // decide in a later pass, before any later hit of their diagonal.
    // Per hit (state per diagonal, flagged = extended to the right):
    //   pending            -> wait
    //   flagged            -> s < last: skip; else last = s, unflag
    //   diff >= window     -> last = s
    //   diff < 3           -> skip
    //   otherwise          -> extension (diagonal pending until it ran)
    // The set and the outcomes travel as lane masks: a vote on one compare is
    // the compare's own result, masks combine in scalar instructions, and a
    // lane's predicate is its bit (inverse_ballot), whereas a vote on a
    // combined predicate costs two VALU instructions (0/1 select and compare).
    __device__ __forceinline__ uint64_t decide_pass(uint64_t am, uint32_t q, int32_t so, uint32_t d,
                                                    int32_t sbase, uint32_t sl)
    {
        // d is a valid index in every lane, active or not
        const bool active = lane_bit(am);
        const uint32_t mk = (uint32_t)(HB_WAVE - lane);
        uint64_t mwin;
        uint32_t st, m = 0;
        if constexpr (G) {
            // markers in the wavefront's LDS table, hashed by diagonal (hits of
            // different diagonals that share a slot are ordered like one
            // diagonal's, which only costs a carry); every hit loads its own
            // state alongside
            const uint32_t slot = d & (kGMarks - 1);
            word_t w0 = 0;
            if (active) {
                w0 = w_ld(d);
                atomicMax(&mkt[slot], mk);
            }
            wave_lds_sync();
            if (active) m = mkt[slot];
            mwin = ballot(m == mk) & am;
            st = (uint32_t)w0 & W::kSM;
        } else {
            word_t old = 0;
            if (active) old = w_max(d, (word_t)mk << W::kMS);
            sync();
            const word_t wv = w_ld(d);
            mwin = ballot((uint32_t)(wv >> W::kMS) == mk) & am;
            // the first hit to reach the word found no marker and holds the
            // state (the max cleared the state bits for the hits after it); a
            // winner that came later gets it from that hit through the word
            const uint64_t mfirst = ballot((uint32_t)(old >> W::kMS) == 0) & am;
            st = (uint32_t)old & W::kSM;
            if (mwin & ~mfirst) {
                if (lane_bit(mfirst & ~mwin)) w_or(d, old);
                sync();
                if (lane_bit(mwin & ~mfirst)) st = (uint32_t)w_ld(d) & W::kSM;
            }
        }
        const bool win = lane_bit(mwin);
        const int32_t sb = so + sbase;
        // s - last on an idle diagonal; negative (below -kP + sb) when the
        // diagonal is pending or flagged
        const int32_t diff = (int32_t)((uint32_t)sb - st);
        // 3 <= diff < window (window > 3)
        const uint64_t mj = ballot((uint32_t)(diff - 3) < (uint32_t)(window - 3)) & mwin;
        const uint64_t mwait = ballot(st == W::kP) & mwin;
        const bool job = lane_bit(mj);
        // the hit becomes last (and clears the flag) on an idle diagonal whose
        // last hit is a window or more behind, or on a flagged one it reached
        const bool set = diff >= window || (uint32_t)(diff + (int32_t)W::kF) < W::kP;
        const uint32_t nst = job ? W::kP : (set ? (uint32_t)sb : st);
        // losers against the winner's new state: wait if it is pending, done
        // if they would skip, else carried
        uint32_t st2;
        if constexpr (G) {
            if (win) {
                w_st(d, (word_t)nst);
                mkt[d & (kGMarks - 1)] = 0;
            }
            // the slot's winner holds the new state when it is on this
            // hit's diagonal; a diagonal without a winner this pass kept its
            // state
            const int wl4 = (int)(((HB_WAVE - m) & 63u) << 2);
            const uint32_t wdg = (uint32_t)__builtin_amdgcn_ds_bpermute(wl4, (int)d);
            const uint32_t wns = (uint32_t)__builtin_amdgcn_ds_bpermute(wl4, (int)nst);
            st2 = wdg == d ? wns : st;
        } else {
            if (win) w_st(d, (word_t)nst);
            sync();
            st2 = (uint32_t)w_ld(d) & W::kSM;
        }
        const int32_t diff2 = (int32_t)((uint32_t)sb - st2);
        const uint64_t mlose = am & ~mwin;
        const uint64_t mb = mwait | (ballot(st2 == W::kP) & mlose);
        const uint64_t mcarry =
            (ballot(diff2 >= 3) | ballot((uint32_t)(diff2 + (int32_t)W::kF) < W::kP)) & mlose;
        const uint32_t tso = (uint32_t)so | (sl << 28);
        if (job)
            jobs[(j_head + n_jobs + lane_rank(mj)) & (kJobCap - 1)] =
                make_uint2(q | ((uint32_t)diff << 24), tso);
        if (lane_bit(mb)) blk[n_blk + lane_rank(mb)] = make_uint2(q, tso);
        n_jobs += (uint32_t)__popcll(mj);
        n_blk += (uint32_t)__popcll(mb);
        sync();
        return mcarry;
    }

    // Decides a set of hits completely (passes until no hit is left).
    __device__ __forceinline__ void decide(uint64_t am, uint32_t q, int32_t so, uint32_t d,
                                           int32_t sbase, uint32_t sl)
    {
        while (am) am = decide_pass(am, q, so, d, sbase, sl);
    }

    // Runs the (up to 64) oldest queued extensions, one per lane, writes HSPs
    // with score >= cutoff and the diagonals' new states.  SL: the jobs'
    // subjects are staged in the ring; else all of them belong to the
    // subject S in global memory (slot 0).
    template <bool SL, typename SP>
    __device__ __forceinline__ void run_batch(SP S)
    {
        const uint32_t nb = min(n_jobs, (uint32_t)HB_WAVE);
        const bool v = (uint32_t)lane < nb;
        ext_result r = {0, 0, 0, 0, 0, 0};
        int32_t k = 0, q = 0, so = 0;
        uint32_t sl = 0, oid = 0;
        if (v) {
            const uint2 jb = jobs[(j_head + (uint32_t)lane) & (kJobCap - 1)];
            q = (int32_t)(jb.x & 0xffffffu);
            so = (int32_t)(jb.y & kSoMask);
            sl = jb.y >> 28;
            const int32_t last_s = so - (int32_t)(jb.x >> 24);
            const uint4 m = sla[sl];
            oid = m.x;
            k = context_of(ctx, n_ctx, q);
            if (SL)
                r = ext_two_hit<true>(mat, ring + m.w, (int32_t)m.y, qry, qlen, last_s + 3, so,
                                      q, ctx[ctx_pad + k]);
            else
                r = ext_two_hit<false>(mat, S, (int32_t)m.y, qry, qlen, last_s + 3, so, q,
                                       ctx[ctx_pad + k]);
            const int32_t b = (int32_t)m.z;
            w_st(diag_of((uint32_t)q, so),
                 (word_t)(r.right ? (W::kF | (uint32_t)(r.s_last - 2 + b)) : (uint32_t)(so + b)));
            atomicAdd(&slb[sl].y, 1u);
        }
        const bool emit = v && r.score >= ctx[2 * ctx_pad + k];
        const uint64_t m = __ballot(emit);
        if (m) {
            uint32_t slot0 = 0;
            if (lane == 0) slot0 = atomicAdd(a->out_count, (uint32_t)__popcll(m));
            slot0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)slot0);
            if (emit) {
                const uint32_t slot = slot0 + lane_rank(m);
                if (slot < a->out_cap) {
                    hb_hsp o;
                    o.oid = oid;
                    o.context = ctx_base + k;
                    o.q_start = r.q_start;
                    o.s_start = r.s_start;
                    o.q_off = q;
                    o.s_off = so;
                    o.len = r.len;
                    o.score = r.score;
                    a->out[slot] = o;
                }
                atomicAdd(&slb[sl].z, 1u);
            }
        }
        j_head = (j_head + nb) & (kJobCap - 1);
        n_jobs -= nb;
        sync();
    }

    // Re-decides the blocked hits in their original order (in place: each
    // round reads its 64 entries before any new entry is written, and new
    // entries never outnumber the ones read).  Every blocked hit waits on a
    // pending diagonal; only the diagonals of the batch just run are no
    // longer pending, so a replay queues at most as many extensions as that
    // batch ran.
    __device__ __forceinline__ void replay()
    {
        const uint32_t nb = n_blk;
        n_blk = 0;
        for (uint32_t b0 = 0; b0 < nb; b0 += HB_WAVE) {
            const uint32_t b = b0 + (uint32_t)lane;
            const bool v = b < nb;
            const uint2 e = v ? blk[b] : make_uint2(0, 0);
            const int32_t so = (int32_t)(e.y & kSoMask);
            const uint32_t sl = e.y >> 28;
            const int32_t sbase = v ? (int32_t)sla[sl].z : 0;
            sync();
            decide(ballot(v), e.x, so, diag_of(e.x, so), sbase, sl);
        }
    }

    // Runs extensions until fewer than 64 are queued and the blocked list
    // has room for another round.
    template <bool SL, typename SP>
    __device__ __forceinline__ void drain_some(SP S)
    {
        while (n_jobs >= (uint32_t)HB_WAVE || (n_blk > kBlkCap - HB_WAVE && n_jobs)) {
            run_batch<SL>(S);
            replay();
        }
    }

    // Runs every queued extension and replays the blocked hits until none is
    // left; every staged subject is then complete and is retired.
    template <bool SL, typename SP>
    __device__ __forceinline__ void finish(SP S)
    {
        while (n_jobs) {
            run_batch<SL>(S);
            replay();
        }
        retire_all();
    }

    // Writes the counters of every staged subject and empties the ring.
    __device__ __forceinline__ void retire_all()
    {
        if (n_live == 0) return;
        uint32_t h = 0, e = 0, hs = 0;
        if ((uint32_t)lane < n_live) {
            const uint32_t sl = (s_tail + (uint32_t)lane) & (HB_SLOTS - 1);
            const uint4 A = sla[sl], B = slb[sl];
            h = B.x;
            e = B.y;
            hs = B.z;
            if (a->stats) {
                a->stats[3 * (size_t)A.x + 0] = h;
                a->stats[3 * (size_t)A.x + 1] = e;
                a->stats[3 * (size_t)A.x + 2] = hs;
            }
        }
        tot_hits += h;
        tot_ext += e;
        tot_hsp += hs;
        n_live = 0;
        s_tail = 0;
        ring_head = 0;
        rwrap = false;
        sync();
    }

    // True when no queued extension belongs to the oldest staged subject.
    // Checked after a replay, none of its hits is then blocked either: a
    // blocked hit waits on a diagonal that a queued extension of its own or
    // an older subject holds, and the oldest subject has no older one.
    __device__ __forceinline__ bool tail_done() const
    {
        bool mine = false;
        for (uint32_t i = (uint32_t)lane; i < n_jobs; i += HB_WAVE)
            mine |= (jobs[(j_head + i) & (kJobCap - 1)].y >> 28) == s_tail;
        return __ballot(mine) == 0;
    }

    // Writes the counters of the oldest staged subject and frees its slot
    // and ring space.
    __device__ __forceinline__ void retire_tail()
    {
        const uint32_t sl = s_tail;
        const uint4 A = sla[sl], B = slb[sl];
        // wave-uniform, so that the ring state stays uniform
        const uint32_t t0 = (uint32_t)__builtin_amdgcn_readfirstlane((int)A.w);
        if (lane == 0) {
            if (a->stats) {
                a->stats[3 * (size_t)A.x + 0] = B.x;
                a->stats[3 * (size_t)A.x + 1] = B.y;
                a->stats[3 * (size_t)A.x + 2] = B.z;
            }
            tot_hits += B.x;
            tot_ext += B.y;
            tot_hsp += B.z;
        }
        s_tail = (sl + 1) & (HB_SLOTS - 1);
        n_live -= 1;
        if (n_live == 0) {   // an empty ring starts over at slot 0, offset 0
            s_tail = 0;
            ring_head = 0;
            rwrap = false;
        } else if (rwrap) {
            // the first subject past the wrap starts at 0, below every
            // offset before it
            const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)sla[s_tail].w);
            if (t < t0) rwrap = false;
        }
    }

    // Returns the ring offset for a new subject of span bytes (span <= the
    // ring size) once a slot and the space are free: retires the oldest
    // subjects whose extensions all ran, and runs the oldest queued
    // extensions while the oldest subject still has some.
    __device__ __forceinline__ uint32_t make_room(uint32_t span)
    {
        for (;;) {
            while (n_live && tail_done()) retire_tail();
            if (n_live == 0) return 0;
            if (n_live < HB_SLOTS) {
                const uint32_t t = (uint32_t)__builtin_amdgcn_readfirstlane((int)sla[s_tail].w);
                if (!rwrap) {
                    if (ring_head + span <= a->ring) return ring_head;
                    if (span <= t) {
                        rwrap = true;
                        return 0;
                    }
                } else if (ring_head + span <= t) {
                    return ring_head;
                }
            }
            run_batch<true>(ring);
            replay();
        }
    }

    // Moves the hits that lost a pass (mask ml) to lanes 0..n-1 in lane
    // order; returns n.  Their values travel in cq/cs/cd: each loser pushes
    // them to the lane of its rank (ds_permute, every lane active); the
    // other lanes push to lane 63, which no loser needs (a pass has a
    // winner, so n <= 63).
    __device__ __forceinline__ uint32_t carry_losers(uint64_t ml, uint32_t q, int32_t so, uint32_t d,
                                                     uint32_t &cq, int32_t &cs, uint32_t &cd)
    {
        const uint32_t n = (uint32_t)__popcll(ml);
        if (n) {
            const int a4 = lane_bit(ml) ? (int)(lane_rank(ml) << 2) : 63 * 4;
            cq = (uint32_t)__builtin_amdgcn_ds_permute(a4, (int)q);
            cs = __builtin_amdgcn_ds_permute(a4, so);
            cd = (uint32_t)__builtin_amdgcn_ds_permute(a4, (int)d);
        }
        return n;
    }

    // One 64-position window of a subject: lane i holds the cell {start,
    // count} of the 3-mer at s0 + i; excl is the exclusive prefix of the
    // counts, or ~0 where the count is 0 (the window's hits in (s, q) order
    // are flat indices 0 .. H-1; hit f of lane i's cell is entry f + eb).
    struct Win {
        uint32_t excl;
        int32_t eb;
        uint32_t H;
        int32_t s0, s0m;   // first position, and it mod D
    };

    // The {start, count} cell of the word at position s0 + lane (0 past the
    // subject's last word).  Every cell is present (count 0 for words
    // without neighbours): a presence-bit test first saves no load for the
    // queries this stage sees and costs an LDS read.
    template <typename SP>
    __device__ __forceinline__ uint64_t cell_at(SP S, int32_t slen, int32_t s0) const
    {
        const int32_t s = s0 + lane;
        uint64_t c2 = 0;
        if (s + 3 <= slen) {
            const uint32_t idx = ((uint32_t)S[s] << 10) | ((uint32_t)S[s + 1] << 5) | S[s + 2];
            c2 = cells[idx];
        }
        return c2;
    }

    // The window at s0 from its cells.  (Loading them a window ahead made
    // the compiler wait on that load where it reused its registers, which
    // also drained the round-ahead entry loads issued after it.)
    __device__ __forceinline__ void make_window(uint64_t c2, int32_t s0, int32_t s0m, Win &w) const
    {
        const uint32_t cst = (uint32_t)c2, cnt = (uint32_t)(c2 >> 32);
        const uint32_t incl = wave_incl_add(cnt);
        const uint32_t excl = incl - cnt;
        w.excl = cnt ? excl : ~0u;
        w.eb = (int32_t)cst - (int32_t)excl;
        w.H = lane63(incl);
        w.s0 = s0;
        w.s0m = s0m;
    }

    // Scans one subject (slot sl, diagonal offset sbase).  Hits are taken in
    // (s, q) order, 64 per round.  The next hits are staged one round ahead
    // in hit order (lane j holds the j-th hit after the last one taken:
    // the rest of the current window and, when that is not enough, the
    // start of the next one), so that their lookup-entry loads run while
    // the current round decides; a round moves the staged hits it takes to
    // its lanes c, c + 1, ... behind the hits it carries.  A hit that lost
    // its diagonal to a lower lane in a round is carried to the front lanes
    // of the next round (ahead of every newer hit, so each diagonal still
    // sees its hits in subject order), and a round makes a single decision
    // pass.  Returns the subject's lookup hits.
    template <bool SL, typename SP>
    __device__ __forceinline__ uint32_t subject(SP S, int32_t slen, uint32_t sl, int32_t sbase)
    {
        uint32_t hits = 0;
        Win A, B;
        A.H = 0;
        A.excl = ~0u;
        A.eb = 0;
        A.s0 = A.s0m = 0;
        B = A;
        bool hasB = false;        // B holds the window after A
        int32_t nxt = 0, nxtm = 0;   // the next window to load, and it mod D
        uint32_t r = 0;           // hits of A already taken
        uint32_t c = 0;           // carried hits, in lanes 0..c-1
        uint32_t cq = 0, cd = 0;
        int32_t cs_ = 0;
        // source of the last hit taken, no larger than any segment start of
        // the next staging (the inclusive max-scan relies on it)
        uint32_t carry = 0;
        // staged hits: n_st of them (n_a from A), lane j: entry (load in
        // flight), position, position mod D
        uint32_t n_st = 0, n_a = 0, eq = 0;
        int32_t eso = 0, esm = 0;
        auto stage = [&]() {
            // A: a window with hits left, unless the subject has no more
            while (r == A.H && (hasB || nxt + 3 <= slen)) {
                if (hasB) {
                    A = B;
                    hasB = false;
                } else {
                    make_window(cell_at(S, slen, nxt), nxt, nxtm, A);
                    hits += A.H;
                    nxt += HB_WAVE;
                    nxtm += HB_WAVE;
                    if (nxtm >= D) nxtm -= D;
                }
                r = 0;
                // the last hit taken was in the old window: the new window's
                // first hit starts a segment
                carry = 0;
            }
            if (!hasB && A.H - r < (uint32_t)HB_WAVE && nxt + 3 <= slen) {
                make_window(cell_at(S, slen, nxt), nxt, nxtm, B);
                hits += B.H;
                hasB = true;
                nxt += HB_WAVE;
                nxtm += HB_WAVE;
                if (nxtm >= D) nxtm -= D;
            }
            const uint32_t t1 = min((uint32_t)HB_WAVE, A.H - r);
            const uint32_t t2 = hasB ? min((uint32_t)HB_WAVE - t1, B.H) : 0u;
            n_a = t1;
            n_st = t1 + t2;
            if (n_st == 0) return;
            // A's hits r .. r + t1 - 1 to lanes 0 .. t1 - 1, then B's 0 ..
            // t2 - 1; segment starts are tagged with their source position
            // + 1 (B's positions count from 64)
            if (++rgen == (1u << 24)) {
                seg[lane] = 0;
                rgen = 1;
            }
            const uint32_t tag = rgen << 8;
            const uint32_t ea = A.excl - r;
            if (ea < t1) seg[ea] = tag | ((uint32_t)lane + 1);
            if (t2 && B.excl < t2) seg[B.excl + t1] = tag | ((uint32_t)lane + 65);
            sync();
            const uint32_t sv = seg[lane] - tag;
            // lanes before the first segment start continue the last taken
            // hit's segment (carry); the source of each lane is the largest
            // start at or before it
            uint32_t v = sv < 256u ? sv : 0u;
            if (lane == 0 && v == 0) v = carry + 1;
            v = wave_incl_max(v);
            const uint32_t p = v - 1;
            const int p4 = (int)((p & 63u) << 2);
            int32_t eb = __builtin_amdgcn_ds_bpermute(p4, A.eb);
            if (t2) {   // B's position p - 64 (its lane p - 64)
                const int32_t ebb = __builtin_amdgcn_ds_bpermute(p4, B.eb);
                eb = p >= 64 ? ebb : eb;
            }
            // (s0 + p) mod D, p < 128 <= 2 D
            int32_t smp = A.s0m + (int32_t)p;
            if (smp >= D) smp -= D;
            if (smp >= D) smp -= D;
            const uint32_t flat = (uint32_t)lane >= t1 ? (uint32_t)lane - t1 : (uint32_t)lane + r;
            eq = 0;
            if ((uint32_t)lane < n_st) eq = entries[(uint32_t)((int32_t)flat + eb)];
            eso = A.s0 + (int32_t)p;
            esm = smp;
            sync();
        };
        stage();
        for (;;) {
            if (n_st + c == 0) break;
            drain_some<SL>(S);
            // this round takes the first m staged hits into lanes c .. c + m - 1
            const uint32_t m = min(n_st, (uint32_t)HB_WAVE - c);
            const int src4 = (int)((((uint32_t)lane - c) & 63u) << 2);
            const uint32_t qn = (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)eq);
            const int32_t son = __builtin_amdgcn_ds_bpermute(src4, eso);
            const int32_t smp = __builtin_amdgcn_ds_bpermute(src4, esm);
            if (m) {
                // the source of the last hit taken, in A's positions (B's
                // count from 64), then the cursor moves past it
                carry = (uint32_t)(__builtin_amdgcn_readlane(eso, (int)(m - 1)) - A.s0);
                if (m <= n_a) {
                    r += m;
                } else {
                    A = B;
                    hasB = false;
                    r = m - n_a;
                    carry -= 64;
                }
            }
            stage();
            int32_t dn = (int32_t)qn - smp;
            if (dn < 0) dn += D;
            const bool isc = (uint32_t)lane < c;
            const uint32_t q = isc ? cq : qn;
            const int32_t so = isc ? cs_ : son;
            const uint32_t d = isc ? cd : (uint32_t)dn;
            // the carried hits and the new ones: lanes 0 .. c + m - 1
            const uint64_t am = c + m >= (uint32_t)HB_WAVE ? ~0ull : (1ull << (c + m)) - 1;
            c = carry_losers(decide_pass(am, q, so, d, sbase, sl), q, so, d, cq, cs_, cd);
        }
        return hits;
    }

    // Scans the subject of processing-order record m.  A subject that fits
    // the ring is staged there (16-byte loads and stores) and its queued
    // extensions wait until 64 are queued, or until its ring space or slot
    // is needed while it is the oldest, so that extension batches stay full;
    // a longer one is scanned from global memory after every staged subject
    // is complete.
    __device__ __forceinline__ void scan_subject(const uint4 m)
    {
        const uint32_t oid = m.x;
        const int32_t slen = (int32_t)m.y;
        const uint8_t *S = a->res + ((uint64_t)m.z | ((uint64_t)m.w << 32));
        const uint32_t span = ((uint32_t)slen + 15u) & ~15u;
        if ((uint32_t)slen <= a->ring) {
            const uint32_t buf = make_room(span);
            const uint32_t sl = (s_tail + n_live) & (HB_SLOTS - 1);
            const uint4 *src = reinterpret_cast<const uint4 *>(S);
            uint4 *dst = reinterpret_cast<uint4 *>(ring + buf);
            for (uint32_t i = (uint32_t)lane; i < span / 16; i += HB_WAVE) dst[i] = src[i];
            if (lane == 0) {
                sla[sl] = make_uint4(oid, (uint32_t)slen, (uint32_t)base, buf);
                slb[sl] = make_uint4(0, 0, 0, 0);
            }
            sync();
            ring_head = buf + span;
            n_live += 1;
            const uint32_t hits = subject<true>(ring + buf, slen, sl, base);
            if (lane == 0) slb[sl].x = hits;
            sync();
        } else {
            finish<true>(ring);
            if (lane == 0) {
                sla[0] = make_uint4(oid, (uint32_t)slen, (uint32_t)base, 0);
                slb[0] = make_uint4(0, 0, 0, 0);
            }
            sync();
            n_live = 1;
            const uint32_t hits = subject<false>(S, slen, 0, base);
            if (lane == 0) slb[0].x = hits;
            sync();
            finish<false>(S);
        }
        if (base + slen + window >= a->state_limit) {
            finish<true>(ring);
            clear_words();
        } else {
            base += slen + window;
        }
    }

    __device__ __forceinline__ void flush_totals(uint32_t g)
    {
        if (a->group_totals && (tot_hits | tot_ext | tot_hsp)) {
            atomicAdd(&a->group_totals[3 * (size_t)g + 0], (unsigned long long)tot_hits);
            atomicAdd(&a->group_totals[3 * (size_t)g + 1], (unsigned long long)tot_ext);
            atomicAdd(&a->group_totals[3 * (size_t)g + 2], (unsigned long long)tot_hsp);
        }
        tot_hits = tot_ext = tot_hsp = 0;
    }
};

// Copies group g's tables into the workgroup's LDS (all threads).
__device__ __forceinline__ void load_group(const hb_scan_args &a, const hb_group_desc &G,
                                           int32_t *sm_ctx, uint8_t *sm_q)
{
    for (int i = threadIdx.x; i < G.qlen_pad / 4; i += blockDim.x)
        reinterpret_cast<uint32_t *>(sm_q)[i] = reinterpret_cast<const uint32_t *>(G.query)[i];
    for (int i = threadIdx.x; i < G.n_ctx; i += blockDim.x) {
        sm_ctx[i] = G.ctx[i];
        sm_ctx[a.ctx_pad_max + i] = G.ctx[G.n_ctx + i];
        sm_ctx[2 * a.ctx_pad_max + i] = G.ctx[2 * G.n_ctx + i];
    }
}

template <bool GS>
__device__ __forceinline__ void bind_group(WaveScan<GS> &w, const hb_group_desc &G)
{
    w.cells = (const HB_GLOBAL uint64_t *)G.cells;
    w.entries = (const HB_GLOBAL uint32_t *)G.entries;
    w.n_entries = G.n_entries > 0 ? G.n_entries : 1u;
    w.qlen = G.qlen;
    w.n_ctx = G.n_ctx;
    w.ctx_base = G.ctx_base;
    w.D = G.diag_len;
    w.dmul = (uint32_t)(0xffffffffu / (uint32_t)G.diag_len) + 1u;
}

// MULTI = false: one group; each wavefront pops subject tiles on its own.
// MULTI = true: many groups; the workgroup pops (group, tile) items and its
// wavefronts share the tile's subjects, so all of them use one group's LDS
// tables at a time.
// GSTATE: the diagonal words live in global memory (a.gstate, diag_len_max
// per wavefront of the grid) instead of LDS, for queries (groups) whose
// state does not fit LDS or would leave it too few wavefronts, and for
// databases whose subjects are too long for the LDS word's positions.
// Every scan kernel keeps a register budget for four wavefronts per SIMD
// (128 VGPRs), the most a 1000-residue query's LDS leaves room for: at five
// (96 VGPRs) the many-group kernel spills to scratch, and 20 wavefronts per
// CU measured slower than 16 where the LDS would allow them.
template <bool MULTI, bool GSTATE>
__device__ __forceinline__ void scan_body(const hb_scan_args &a);

template <bool MULTI, bool GSTATE>
__global__ __launch_bounds__(HB_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void hb_scan_kernel(hb_scan_args a)
{
    scan_body<MULTI, GSTATE>(a);
}
template <bool MULTI>
__global__ __launch_bounds__(HB_BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void hb_scan_kernel_g(hb_scan_args a)
{
    scan_body<MULTI, true>(a);
}

template <bool MULTI, bool GSTATE>
__device__ __forceinline__ void scan_body(const hb_scan_args &a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ uint32_t s_item, s_next;
    const int lane = threadIdx.x & (HB_WAVE - 1);
    const int wid = threadIdx.x / HB_WAVE;
    int8_t *sm_mat = reinterpret_cast<int8_t *>(lds);
    int32_t *sm_ctx = reinterpret_cast<int32_t *>(lds + 1024);
    uint8_t *sm_q = reinterpret_cast<uint8_t *>(sm_ctx + 3 * a.ctx_pad_max);
    unsigned char *wb = sm_q + a.qlen_pad_max + (size_t)wid * a.wave_lds;
    const int32_t Dm = a.diag_len_max;

    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        reinterpret_cast<uint32_t *>(sm_mat)[i] = reinterpret_cast<const uint32_t *>(a.mat32)[i];

    WaveScan<GSTATE> w;
    w.a = &a;
    w.mat = sm_mat;
    w.qry = sm_q;
    w.ctx = sm_ctx;
    w.ctx_pad = a.ctx_pad_max;
    if constexpr (GSTATE) {
        w.wd = a.gstate + ((size_t)blockIdx.x * (blockDim.x / HB_WAVE) + (size_t)wid) * (size_t)Dm;
        w.mkt = reinterpret_cast<uint32_t *>(wb);
        for (int i = lane; i < (int)kGMarks; i += HB_WAVE) w.mkt[i] = 0;
        w.seg = w.mkt + kGMarks;
    } else {
        w.wd = reinterpret_cast<uint32_t *>(wb);
        w.mkt = nullptr;
        w.seg = reinterpret_cast<uint32_t *>(wb + 4 * (size_t)Dm);   // Dm: multiple of 64
    }
    w.jobs = reinterpret_cast<uint2 *>(w.seg + HB_WAVE);
    w.blk = w.jobs + kJobCap;
    w.sla = reinterpret_cast<uint4 *>(w.blk + kBlkCap);
    w.slb = w.sla + HB_SLOTS;
    w.ring = reinterpret_cast<uint8_t *>(w.slb + HB_SLOTS);
    w.lane = lane;
    w.window = a.window;
    w.rgen = 0;
    w.j_head = 0;
    w.n_jobs = 0;
    w.n_blk = 0;
    w.n_live = 0;
    w.s_tail = 0;
    w.ring_head = 0;
    w.rwrap = false;
    w.tot_hits = w.tot_ext = w.tot_hsp = 0;
    w.seg[lane] = 0;
    const uint32_t n_tiles = a.tile_end - a.tile_begin;

    if (!MULTI) {
        const hb_group_desc G = a.groups[0];
        load_group(a, G, sm_ctx, sm_q);
        bind_group(w, G);
        w.clear_words();
        __syncthreads();
        for (;;) {
            uint32_t t = 0;
            if (lane == 0) t = atomicAdd(a.work_counter, 1u);
            t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
            if (t >= n_tiles) break;
            t += a.tile_begin;
            const uint32_t i_end = a.tiles[t + 1];
            for (uint32_t it = a.tiles[t]; it < i_end; ++it) w.scan_subject(a.subj[it]);
        }
        w.template finish<true>(w.ring);
        w.flush_totals(0);
    } else {
        uint32_t cur = 0xffffffffu;
        const uint32_t tpi = a.tiles_per_item;
        const uint32_t n_chunks = (n_tiles + tpi - 1) / tpi;
        const uint32_t n_items = n_chunks * a.n_groups;
        for (;;) {
            if (threadIdx.x == 0) s_item = atomicAdd(a.work_counter, 1u);
            __syncthreads();
            const uint32_t item = s_item;
            if (item >= n_items) break;
            const uint32_t g = item / n_chunks;
            const uint32_t t0 = a.tile_begin + (item % n_chunks) * tpi;
            const uint32_t t1 = min(t0 + tpi, a.tile_end);
            if (g != cur) {
                if (cur != 0xffffffffu) {
                    w.template finish<true>(w.ring);   // still with the old group's tables
                    w.flush_totals(cur);
                }
                __syncthreads();
                const hb_group_desc G = a.groups[g];
                load_group(a, G, sm_ctx, sm_q);
                bind_group(w, G);
                w.clear_words();
                cur = g;
                __syncthreads();
            }
            // the chunk's subjects are taken one at a time by whichever
            // wavefront is free (LDS counter), so the workgroup's wavefronts
            // reach the next barrier at most one subject apart
            if (threadIdx.x == 0) s_next = a.tiles[t0];
            __syncthreads();
            const uint32_t i_end = a.tiles[t1];
            for (;;) {
                uint32_t it = 0;
                if (lane == 0) it = atomicAdd(&s_next, 1u);
                it = (uint32_t)__builtin_amdgcn_readfirstlane((int)it);
                if (it >= i_end) break;
                w.scan_subject(a.subj[it]);
            }
            __syncthreads();
        }
        if (cur != 0xffffffffu) {
            w.template finish<true>(w.ring);
            w.flush_totals(cur);
        }
    }
}

// ---------------------------------------------------------------- results
// Device-side result layout for the plain form: per-subject counters become
// hb_subject_stats with the index of the subject's first record (exclusive
// prefix of HSP counts over oids), records are scattered to their subject's
// slot and each subject's records are ordered by (s_off, q_off) -- the
// order in which the CPU word finder discovers them.
constexpr int kFinThreads = 1024;
constexpr int kFinPer = 4;
constexpr uint32_t kFinSpan = kFinThreads * kFinPer;   // subjects per block

// Block-wide exclusive scan of v (kFinThreads threads); returns the block total in *total.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sm, uint32_t *total)
{
    const int lane = threadIdx.x & (HB_WAVE - 1), wid = threadIdx.x / HB_WAVE;
    const uint32_t incl = wave_incl_add(v);
    if (lane == HB_WAVE - 1) sm[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        const uint32_t w = lane < kFinThreads / HB_WAVE ? sm[lane] : 0u;
        const uint32_t wi = wave_incl_add(w);
        if (lane < kFinThreads / HB_WAVE) sm[lane] = wi - w;
        if (lane == HB_WAVE - 1) sm[kFinThreads / HB_WAVE] = wi;
    }
    __syncthreads();
    const uint32_t r = sm[wid] + incl - v;
    *total = sm[kFinThreads / HB_WAVE];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kFinThreads) void fin_sum_kernel(const uint32_t *st3, uint32_t ns,
                                                              uint32_t *part)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    uint32_t v = 0;
    for (int k = 0; k < kFinPer; ++k) {
        const uint32_t i = blockIdx.x * kFinSpan + (uint32_t)k * kFinThreads + threadIdx.x;
        if (i < ns) v += st3[3 * (size_t)i + 2];
    }
    uint32_t total = 0;
    block_excl_scan(v, sm, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
}

// One block: exclusive scan of the nb block totals in place; the grand
// total goes to *total_out when it is given.
__global__ __launch_bounds__(kFinThreads) void fin_part_kernel(uint32_t *part, uint32_t nb,
                                                               uint32_t *total_out)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += kFinThreads) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t v = i < nb ? part[i] : 0u;
        uint32_t total = 0;
        const uint32_t e = block_excl_scan(v, sm, &total);
        if (i < nb) part[i] = carry + e;
        carry += total;
    }
    if (total_out && threadIdx.x == 0) *total_out = carry;
}

// Compact layout, first pass: per block, the HSP count and the number of
// subjects with HSPs; the launch's lookup hits and extensions are summed
// into tot[0], tot[1] (one 64-bit atomic per wavefront).
__global__ __launch_bounds__(kFinThreads) void fin_sum2_kernel(const uint32_t *st3, uint32_t ns,
                                                               uint32_t *part, uint32_t *part_s,
                                                               unsigned long long *tot)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    uint32_t v = 0, nz = 0;
    unsigned long long lh = 0, ex = 0;
    for (int k = 0; k < kFinPer; ++k) {
        const uint32_t i = blockIdx.x * kFinSpan + (uint32_t)k * kFinThreads + threadIdx.x;
        if (i < ns) {
            const uint32_t c = st3[3 * (size_t)i + 2];
            v += c;
            nz += c != 0;
            lh += st3[3 * (size_t)i];
            ex += st3[3 * (size_t)i + 1];
        }
    }
    uint32_t total = 0;
    block_excl_scan(v, sm, &total);
    if (threadIdx.x == 0) part[blockIdx.x] = total;
    block_excl_scan(nz, sm, &total);
    if (threadIdx.x == 0) part_s[blockIdx.x] = total;
    for (int o = HB_WAVE / 2; o > 0; o >>= 1) {
        lh += __shfl_xor(lh, o, HB_WAVE);
        ex += __shfl_xor(ex, o, HB_WAVE);
    }
    // the block's wavefront sums, then one atomic pair per block
    __shared__ unsigned long long s64[2 * (kFinThreads / HB_WAVE)];
    const int wid = threadIdx.x / HB_WAVE;
    if ((threadIdx.x & (HB_WAVE - 1)) == 0) {
        s64[2 * wid] = lh;
        s64[2 * wid + 1] = ex;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long a = 0, b = 0;
        for (int w = 0; w < kFinThreads / HB_WAVE; ++w) {
            a += s64[2 * w];
            b += s64[2 * w + 1];
        }
        atomicAdd(&tot[0], a);
        atomicAdd(&tot[1], b);
    }
}

// Compact layout, second pass: every subject with HSPs gets its entry at
// its rank among them (so in oid order), and every subject its first
// record index in cursor[] for the scatter.
__global__ __launch_bounds__(kFinThreads) void fin_compact_kernel(const uint32_t *st3, uint32_t ns,
                                                                  const uint32_t *part,
                                                                  const uint32_t *part_s,
                                                                  hb_subject_hits *out,
                                                                  uint32_t *cursor)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    const uint32_t i0 = blockIdx.x * kFinSpan + (uint32_t)threadIdx.x * kFinPer;
    uint32_t cnt[kFinPer];
    uint32_t v = 0, nz = 0;
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        cnt[k] = i0 + k < ns ? st3[3 * (size_t)(i0 + k) + 2] : 0u;
        v += cnt[k];
        nz += cnt[k] != 0;
    }
    uint32_t total = 0;
    uint32_t first = part[blockIdx.x] + block_excl_scan(v, sm, &total);
    uint32_t rank = part_s[blockIdx.x] + block_excl_scan(nz, sm, &total);
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        const uint32_t i = i0 + k;
        if (i < ns) {
            cursor[i] = first;
            if (cnt[k]) {
                hb_subject_hits o;
                o.oid = i;
                o.n_hsp = cnt[k];
                o.lookup_hits = st3[3 * (size_t)i];
                o.extensions = st3[3 * (size_t)i + 1];
                o.first = first;
                out[rank++] = o;
            }
        }
        first += cnt[k];
    }
}

__global__ __launch_bounds__(kFinThreads) void fin_write_kernel(const uint32_t *st3, uint32_t ns,
                                                                const uint32_t *part,
                                                                hb_subject_stats *out,
                                                                uint32_t *cursor)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    const uint32_t i0 = blockIdx.x * kFinSpan + (uint32_t)threadIdx.x * kFinPer;
    uint32_t cnt[kFinPer];
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        cnt[k] = i0 + k < ns ? st3[3 * (size_t)(i0 + k) + 2] : 0u;
        v += cnt[k];
    }
    uint32_t total = 0;
    uint32_t first = part[blockIdx.x] + block_excl_scan(v, sm, &total);
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        const uint32_t i = i0 + k;
        if (i < ns) {
            hb_subject_stats o;
            o.lookup_hits = st3[3 * (size_t)i];
            o.extensions = st3[3 * (size_t)i + 1];
            o.n_hsp = cnt[k];
            o.first = cnt[k] ? (int64_t)first : -1;
            out[i] = o;
            cursor[i] = first;
        }
        first += cnt[k];
    }
}

// First-record index of each subject from its count (no stats written).
__global__ __launch_bounds__(kFinThreads) void fin_first_kernel(const uint32_t *st3, uint32_t ns,
                                                                const uint32_t *part,
                                                                uint32_t *first_out)
{
    __shared__ uint32_t sm[kFinThreads / HB_WAVE + 1];
    const uint32_t i0 = blockIdx.x * kFinSpan + (uint32_t)threadIdx.x * kFinPer;
    uint32_t cnt[kFinPer];
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        cnt[k] = i0 + k < ns ? st3[3 * (size_t)(i0 + k) + 2] : 0u;
        v += cnt[k];
    }
    uint32_t total = 0;
    uint32_t first = part[blockIdx.x] + block_excl_scan(v, sm, &total);
#pragma unroll
    for (int k = 0; k < kFinPer; ++k) {
        if (i0 + k < ns) first_out[i0 + k] = first;
        first += cnt[k];
    }
}

__global__ void fin_scatter_kernel(const hb_hsp *in, uint32_t n, uint32_t *cursor, hb_hsp *out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const hb_hsp h = in[i];
        out[atomicAdd(&cursor[h.oid], 1u)] = h;
    }
}

// Insertion sort of one subject's n records at recs + f by (s_off, q_off);
// a subject with more than HB_FIN_SORT_MAX records is listed for the host.
__device__ void sort_subject(uint32_t oid, uint32_t n, uint64_t f, hb_hsp *recs, hb_fin_big *big,
                             uint32_t big_cap, uint32_t *n_big)
{
    if (n < 2) return;
    if (n > HB_FIN_SORT_MAX) {
        const uint32_t k = atomicAdd(n_big, 1u);
        if (k < big_cap) big[k] = hb_fin_big{oid, n, f};
        return;
    }
    hb_hsp *r = recs + f;
    for (uint32_t a = 1; a < n; ++a) {
        const hb_hsp x = r[a];
        uint32_t b = a;
        while (b > 0 && (r[b - 1].s_off > x.s_off ||
                         (r[b - 1].s_off == x.s_off && r[b - 1].q_off > x.q_off))) {
            r[b] = r[b - 1];
            --b;
        }
        r[b] = x;
    }
}

// One thread per subject (dense layout).
__global__ void fin_sort_kernel(const hb_subject_stats *st, uint32_t ns, hb_hsp *recs,
                                hb_fin_big *big, uint32_t big_cap, uint32_t *n_big)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ns) sort_subject(i, st[i].n_hsp, (uint64_t)st[i].first, recs, big, big_cap, n_big);
}

// One thread per subject with HSPs (compact layout; their number is *n_subj).
__global__ void fin_sort_compact_kernel(const hb_subject_hits *sh, const uint32_t *n_subj,
                                        hb_hsp *recs, hb_fin_big *big, uint32_t big_cap,
                                        uint32_t *n_big)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < *n_subj) sort_subject(sh[i].oid, sh[i].n_hsp, sh[i].first, recs, big, big_cap, n_big);
}

// Plain 16-byte-per-lane copy: the known-size reference kernel for
// calibrating HBM byte counters and the achievable copy bandwidth.
__global__ __launch_bounds__(256) void copy_kernel(uint4 *dst, const uint4 *src, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

__global__ void fill_u32_kernel(uint32_t *dst, uint32_t value, uint32_t n)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        dst[i] = value;
}

}  // namespace

hipError_t hb_launch_copy(void *dst, const void *src, uint64_t bytes, hipStream_t st)
{
    const uint64_t n16 = bytes / 16;
    uint64_t blocks = (n16 + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                       reinterpret_cast<uint4 *>(dst), reinterpret_cast<const uint4 *>(src), n16);
    return hipGetLastError();
}

hipError_t hb_launch_finalize(const uint32_t *st3, uint32_t ns, const hb_hsp *recs, uint32_t n,
                              hb_subject_stats *stats_out, hb_hsp *recs_out, uint32_t *scratch,
                              hb_fin_big *big, uint32_t big_cap, uint32_t *n_big, hipStream_t st)
{
    if (ns == 0) return hipSuccess;
    const uint32_t nb = (ns + kFinSpan - 1) / kFinSpan;
    uint32_t *part = scratch;        // nb
    uint32_t *cursor = scratch + nb; // ns
    hipLaunchKernelGGL(fin_sum_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, part);
    hipLaunchKernelGGL(fin_part_kernel, dim3(1), dim3(kFinThreads), 0, st, part, nb, nullptr);
    hipLaunchKernelGGL(fin_write_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, part,
                       stats_out, cursor);
    if (n) {
        hipLaunchKernelGGL(fin_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, recs, n,
                           cursor, recs_out);
        hipLaunchKernelGGL(fin_sort_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, stats_out,
                           ns, recs_out, big, big_cap, n_big);
    }
    return hipGetLastError();
}

hipError_t hb_launch_finalize_compact(const uint32_t *st3, uint32_t ns, const hb_hsp *recs,
                                      uint32_t n, hb_subject_hits *subj_out, hb_hsp *recs_out,
                                      uint32_t *scratch, unsigned long long *tot, hb_fin_big *big,
                                      uint32_t big_cap, uint32_t *n_big, hipStream_t st)
{
    if (ns == 0) return hipSuccess;
    const uint32_t nb = (ns + kFinSpan - 1) / kFinSpan;
    uint32_t *part = scratch;              // nb
    uint32_t *part_s = scratch + nb;       // nb
    uint32_t *n_subj = scratch + 2 * nb;   // 1
    uint32_t *cursor = n_subj + 1;         // ns
    hipLaunchKernelGGL(fin_sum2_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, part, part_s,
                       tot);
    hipLaunchKernelGGL(fin_part_kernel, dim3(1), dim3(kFinThreads), 0, st, part, nb, nullptr);
    hipLaunchKernelGGL(fin_part_kernel, dim3(1), dim3(kFinThreads), 0, st, part_s, nb, n_subj);
    hipLaunchKernelGGL(fin_compact_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, part,
                       part_s, subj_out, cursor);
    if (n) {
        hipLaunchKernelGGL(fin_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, st, recs, n,
                           cursor, recs_out);
        const uint32_t grid = ((ns < n ? ns : n) + 255) / 256;
        hipLaunchKernelGGL(fin_sort_compact_kernel, dim3(grid), dim3(256), 0, st, subj_out, n_subj,
                           recs_out, big, big_cap, n_big);
    }
    return hipGetLastError();
}

hipError_t hb_launch_first_index(const uint32_t *st3, uint32_t ns, uint32_t *first,
                                 uint32_t *scratch, hipStream_t st)
{
    if (ns == 0) return hipSuccess;
    const uint32_t nb = (ns + kFinSpan - 1) / kFinSpan;
    hipLaunchKernelGGL(fin_sum_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, scratch);
    hipLaunchKernelGGL(fin_part_kernel, dim3(1), dim3(kFinThreads), 0, st, scratch, nb, nullptr);
    hipLaunchKernelGGL(fin_first_kernel, dim3(nb), dim3(kFinThreads), 0, st, st3, ns, scratch,
                       first);
    return hipGetLastError();
}

hipError_t hb_launch_fill_u32(uint32_t *dst, uint32_t value, uint32_t n, hipStream_t st)
{
    const uint32_t blocks = n ? (n + 255) / 256 : 1;
    hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks < 1024 ? blocks : 1024), dim3(256), 0, st, dst,
                       value, n);
    return hipGetLastError();
}

__global__ void gather_subjects_kernel(const uint8_t *res, const uint64_t *offs, const uint4 *subj,
                                       const uint32_t *tiles, uint8_t *out)
{
    const uint32_t t = blockIdx.x;
    for (uint32_t i = tiles[t]; i < tiles[t + 1]; ++i) {
        const uint4 m = subj[i];
        const uint8_t *src = res + offs[m.x];
        uint8_t *dst = out + ((uint64_t)m.z | ((uint64_t)m.w << 32));
        for (uint32_t j = threadIdx.x; j < m.y; j += blockDim.x) dst[j] = src[j];
    }
}

hipError_t hb_launch_gather_subjects(const uint8_t *res, const uint64_t *offs, const uint4 *subj,
                                     const uint32_t *tiles, uint32_t n_tiles, uint8_t *out,
                                     hipStream_t st)
{
    if (n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_subjects_kernel, dim3(n_tiles), dim3(256), 0, st, res, offs, subj,
                       tiles, out);
    return hipGetLastError();
}

hipError_t hb_launch_check_residues(const uint8_t *res, uint64_t n, uint32_t *flag,
                                    hipStream_t st)
{
    if (n == 0) return hipSuccess;
    uint64_t blocks = (n / 4 + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(check_residues_kernel, dim3((unsigned)blocks), dim3(256), 0, st, res, n,
                       flag);
    return hipGetLastError();
}

hipError_t hb_launch_lut(const hb_lut_job *jobs_dev, int n_jobs, int32_t qlen_max,
                         const int8_t *mat32, int32_t threshold, int pass, hipStream_t st)
{
    if (n_jobs <= 0) return hipSuccess;
    const size_t lds = 1024 + (size_t)((qlen_max + 15) & ~15);
    if (pass == 1) {
        hipLaunchKernelGGL(lut_scan_kernel, dim3(1, n_jobs), dim3(1024), 0, st, jobs_dev);
        return hipGetLastError();
    }
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(pass == 0 ? lut_kernel<0>
                                                                                : lut_kernel<2>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (pass == 0)
        hipLaunchKernelGGL(lut_kernel<0>, dim3(HB_NCELLS / 256, n_jobs), dim3(256), lds, st,
                           jobs_dev, mat32, threshold);
    else
        hipLaunchKernelGGL(lut_kernel<2>, dim3(HB_NCELLS / 256, n_jobs), dim3(256), lds, st,
                           jobs_dev, mat32, threshold);
    return hipGetLastError();
}

// Instantiations: (MULTI, GSTATE).
static const void *scan_fn(bool multi, bool gstate)
{
    if (gstate)
        return multi ? reinterpret_cast<const void *>(hb_scan_kernel_g<true>)
                     : reinterpret_cast<const void *>(hb_scan_kernel_g<false>);
    return multi ? reinterpret_cast<const void *>(hb_scan_kernel<true, false>)
                 : reinterpret_cast<const void *>(hb_scan_kernel<false, false>);
}

hipError_t hb_launch_scan(const hb_scan_args &a, bool multi, int n_blocks, int block_threads,
                          size_t lds_bytes, hipStream_t st)
{
    if (block_threads < HB_WAVE || block_threads > HB_BLOCK || block_threads % HB_WAVE)
        return hipErrorInvalidValue;
    if (a.ring % 16 || a.wave_lds % 16 || a.state_limit <= 0) return hipErrorInvalidValue;
    const bool g = a.gstate != nullptr;
    const dim3 grid(n_blocks), block(block_threads);
    if (multi) {
        if (g) hipLaunchKernelGGL(hb_scan_kernel_g<true>, grid, block, lds_bytes, st, a);
        else hipLaunchKernelGGL((hb_scan_kernel<true, false>), grid, block, lds_bytes, st, a);
    } else {
        if (g) hipLaunchKernelGGL(hb_scan_kernel_g<false>, grid, block, lds_bytes, st, a);
        else hipLaunchKernelGGL((hb_scan_kernel<false, false>), grid, block, lds_bytes, st, a);
    }
    return hipGetLastError();
}

hipError_t hb_scan_allow_lds(size_t max_lds)
{
    for (int k = 0; k < 4; ++k) {
        const void *fn = scan_fn((k & 1) != 0, (k & 2) != 0);
        hipFuncAttributes fa;
        hipError_t e = hipFuncGetAttributes(&fa, fn);
        if (e != hipSuccess) return e;
        // the limit covers dynamic LDS only; the kernel's static LDS comes on top
        e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(max_lds - fa.sharedSizeBytes));
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t hb_scan_occupancy(bool multi, bool gstate, size_t lds_bytes, int block_threads,
                             int *blocks_per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, scan_fn(multi, gstate),
                                                        block_threads, lds_bytes);
}
