// hb_kernels.hip -- gfx950 kernels of the seed-extension stage.
//
// Scan kernel (hb_scan_kernel): the BLASTP/BLASTX word scan, two-hit
// diagonal tracking and X-drop ungapped extension of
// h_blast_v1.1/H-BLAST/h_blast_kernel.cu:21-1540, restructured for CDNA4:
//   * one wavefront owns one subject at a time; subjects come from a
//     length-descending tile queue (one atomic per tile);
//   * lane i owns subject position s0+i of a 64-position window; the 3-mer
//     index is tested against the presence vector in LDS and the cell
//     {start, count} fetched from the lookup table (L2 resident);
//   * the window's hits are flattened in (s, q) order into an LDS list by a
//     wavefront prefix sum, then processed 64 at a time;
//   * the per-diagonal state {last_hit + offset | flag} lives in LDS (int32
//     per diagonal, no 15-bit limit).  Hits of one round that fall on the
//     same diagonal are ordered by an LDS atomic-max marker so that every
//     diagonal sees its hits in subject order -- the exact sequential
//     semantics of the CPU two-hit word finder (the reference kernel's
//     deferral, h_blast_kernel.cu:873-942, is not reproduced);
//   * X-drop extensions read the query and the score matrix from LDS.
#include <hip/hip_runtime.h>

#include "hb_internal.h"

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

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) { return lane63(wave_incl_add(v)); }

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
// qflag[q] = residue | 0x80 when the 3-mer at q lies inside one context.
// Cell w = (a<<10)|(b<<5)|c holds q when M[q0][a]+M[q1][b]+M[q2][c] >= T or
// w is the query word itself (an exact word always seeds).
template <bool FILL>
__global__ __launch_bounds__(256) void lut_kernel(const uint8_t *qflag, int32_t qlen,
                                                  const int8_t *mat32, int32_t threshold,
                                                  uint32_t *count, const uint32_t *start,
                                                  uint32_t *entries, uint2 *cells,
                                                  uint32_t *pv)
{
    extern __shared__ __align__(16) unsigned char lds[];
    int8_t *sm = reinterpret_cast<int8_t *>(lds);        // 32x32
    uint8_t *sq = lds + 1024;                              // qlen bytes
    for (int i = threadIdx.x; i < 1024; i += blockDim.x) sm[i] = mat32[i];
    for (int i = threadIdx.x; i < qlen; i += blockDim.x) sq[i] = qflag[i];
    __syncthreads();
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= HB_NCELLS) return;
    const uint32_t a = w >> 10, b = (w >> 5) & 31, c = w & 31;
    uint32_t n = 0;
    if (a < HB_ALPHA && b < HB_ALPHA && c < HB_ALPHA) {
        uint32_t k = FILL ? start[w] : 0;
        for (int32_t q = 0; q + 3 <= qlen; ++q) {
            const uint8_t f0 = sq[q];
            if (!(f0 & 0x80)) continue;
            const uint32_t q0 = f0 & 31, q1 = sq[q + 1] & 31, q2 = sq[q + 2] & 31;
            const int32_t s = sm[(q0 << 5) | a] + sm[(q1 << 5) | b] + sm[(q2 << 5) | c];
            const bool self = (q0 == a) & (q1 == b) & (q2 == c);
            if (s >= threshold || self) {
                if (FILL) entries[k++] = (uint32_t)q;
                ++n;
            }
        }
    }
    if (!FILL) {
        count[w] = n;
    } else {
        cells[w] = make_uint2(start[w], n);
        const uint64_t m = __ballot(n != 0);
        const int lane = threadIdx.x & 63;
        if (lane == 0) pv[w >> 5] = (uint32_t)m;
        if (lane == 32) pv[w >> 5] = (uint32_t)(m >> 32);
    }
}

// Exclusive scan of HB_NCELLS counts in place; start[HB_NCELLS] = total.
__global__ __launch_bounds__(1024) void lut_scan_kernel(uint32_t *cnt, uint32_t *total)
{
    __shared__ uint32_t part[1024];
    const int t = threadIdx.x;
    const int per = HB_NCELLS / 1024;
    uint32_t s = 0;
    for (int i = 0; i < per; ++i) s += cnt[t * per + i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        uint32_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint32_t run = part[t] - s;
    for (int i = 0; i < per; ++i) {
        uint32_t v = cnt[t * per + i];
        cnt[t * per + i] = run;
        run += v;
    }
    if (t == 1023) {
        cnt[HB_NCELLS] = run;
        *total = run;
    }
}

// ---------------------------------------------------------------- extensions
// Left X-drop walk from (s_off, q_off) down to offset 0
// (h_blast_kernel.cu:1329-1540): positions k = 0..n, n = min(s_off, q_off);
// strict '>' keeps the first best; stops once max - score >= X.  Returns the
// best score; *len = positions kept (0 if no prefix improved on maxscore).
// Residues are fetched four steps at a time so that the LDS/L1 latency is
// paid once per four steps.
template <typename SP>
__device__ __forceinline__ int32_t ext_left(const int8_t *m, SP S, const uint8_t *Q,
                                            int32_t s_off, int32_t q_off, int32_t X,
                                            int32_t maxscore, int32_t *len)
{
    const int32_t n = min(s_off, q_off);
    int32_t best = -1, sc = maxscore, k = 0;
    bool done = false;
    for (; k + 4 <= n + 1; k += 4) {
        const int32_t m0 = m[((uint32_t)Q[q_off - k] << 5) | S[s_off - k]];
        const int32_t m1 = m[((uint32_t)Q[q_off - k - 1] << 5) | S[s_off - k - 1]];
        const int32_t m2 = m[((uint32_t)Q[q_off - k - 2] << 5) | S[s_off - k - 2]];
        const int32_t m3 = m[((uint32_t)Q[q_off - k - 3] << 5) | S[s_off - k - 3]];
        sc += m0;
        if (sc > maxscore) { maxscore = sc; best = k; }
        if (maxscore - sc >= X) { done = true; break; }
        sc += m1;
        if (sc > maxscore) { maxscore = sc; best = k + 1; }
        if (maxscore - sc >= X) { done = true; break; }
        sc += m2;
        if (sc > maxscore) { maxscore = sc; best = k + 2; }
        if (maxscore - sc >= X) { done = true; break; }
        sc += m3;
        if (sc > maxscore) { maxscore = sc; best = k + 3; }
        if (maxscore - sc >= X) { done = true; break; }
    }
    if (!done) {
        for (; k <= n; ++k) {
            sc += m[((uint32_t)Q[q_off - k] << 5) | S[s_off - k]];
            if (sc > maxscore) { maxscore = sc; best = k; }
            if (maxscore - sc >= X) break;
        }
    }
    *len = best + 1;
    return maxscore;
}

// Right X-drop walk (h_blast_kernel.cu:1133-1326): positions i = 0..n-1,
// n = min(slen - s_off, qlen - q_off); also stops at score <= 0;
// *s_last = s_off + stopping index (n if the walk ran out).
template <typename SP>
__device__ __forceinline__ int32_t ext_right(const int8_t *m, SP S, int32_t slen, const uint8_t *Q,
                                             int32_t qlen, int32_t s_off, int32_t q_off, int32_t X,
                                             int32_t maxscore, int32_t *len, int32_t *s_last)
{
    const int32_t n = min(slen - s_off, qlen - q_off);
    int32_t best = -1, sc = maxscore, i = 0;
    bool done = false;
    for (; i + 4 <= n; i += 4) {
        const int32_t m0 = m[((uint32_t)Q[q_off + i] << 5) | S[s_off + i]];
        const int32_t m1 = m[((uint32_t)Q[q_off + i + 1] << 5) | S[s_off + i + 1]];
        const int32_t m2 = m[((uint32_t)Q[q_off + i + 2] << 5) | S[s_off + i + 2]];
        const int32_t m3 = m[((uint32_t)Q[q_off + i + 3] << 5) | S[s_off + i + 3]];
        sc += m0;
        if (sc > maxscore) { maxscore = sc; best = i; }
        if (sc <= 0 || maxscore - sc >= X) { done = true; break; }
        sc += m1;
        if (sc > maxscore) { maxscore = sc; best = i + 1; }
        if (sc <= 0 || maxscore - sc >= X) { i += 1; done = true; break; }
        sc += m2;
        if (sc > maxscore) { maxscore = sc; best = i + 2; }
        if (sc <= 0 || maxscore - sc >= X) { i += 2; done = true; break; }
        sc += m3;
        if (sc > maxscore) { maxscore = sc; best = i + 3; }
        if (sc <= 0 || maxscore - sc >= X) { i += 3; done = true; break; }
    }
    if (!done) {
        for (; i < n; ++i) {
            sc += m[((uint32_t)Q[q_off + i] << 5) | S[s_off + i]];
            if (sc > maxscore) { maxscore = sc; best = i; }
            if (sc <= 0 || maxscore - sc >= X) break;
        }
    }
    *len = best + 1;
    *s_last = s_off + i;
    return maxscore;
}

struct ext_result {
    int32_t q_start, s_start, len, score, right, s_last;
};

// Two-hit extension (h_blast_kernel.cu:1069-1130): start just after the
// best-scoring prefix of the second word, extend left from maxscore 0, and
// extend right from the left score only if the left walk reached
// s_left_off (the end of the first hit's word).
template <typename SP>
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
    left_score = ext_left(m, S, Q, s_right_off - 1, q_right_off - 1, X, 0, &left_len);
    if (left_len >= s_right_off - s_left_off) {
        r.right = 1;
        right_score = ext_right(m, S, slen, Q, qlen, s_right_off, q_right_off, X, left_score,
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
// Per-wavefront LDS of the scan:
//   state[D]   int32  diagonal state: bit 31 = extended-right flag,
//                     bit 30 = extension pending, bits 0..29 = last hit + offset
//   marker[D]  uint32 same-round ordering marker (generation << 6 | 63 - lane)
//   seg[64]    uint32 tagged segment starts of the current hit round
//   jobs[kJobCap] uint2 pending two-hit extensions (q | (s - last) << 24, s)
//   blk[kBlkCap]  uint2 hits waiting on a pending diagonal (q, s)
//   sbuf[kSCap]   residues of the current subject (if it fits)
constexpr int kJobCap = HB_JOBCAP;
constexpr int kBlkCap = HB_BLKCAP;
constexpr int kSCap = HB_SCAP;
constexpr uint32_t kFlag = 0x80000000u;
constexpr uint32_t kPend = 0x40000000u;
constexpr uint32_t kVal = 0x3fffffffu;

__device__ __forceinline__ uint32_t lane_rank(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

struct WaveScan {
    const hb_scan_args *a;
    const int8_t *mat;
    const uint8_t *qry;
    const int32_t *ctx;   // LDS: ctx_off[ctx_pad], xdrop[ctx_pad], cutoff[ctx_pad]
    const uint32_t *pv;
    int32_t *state;
    uint32_t *marker;
    uint32_t *seg;
    uint2 *jobs;
    uint2 *blk;
    uint8_t *sbuf;
    int lane;
    int32_t D;
    int32_t window;
    int32_t base;      // diagonal offset: stale entries read as "no recent hit"
    uint32_t gen;      // marker generation
    uint32_t rgen;     // segment-tag generation
    uint32_t n_jobs;   // wave-uniform
    uint32_t n_blk;    // wave-uniform
    uint32_t n_ext;    // per lane
    uint32_t n_hsp;    // per lane

    __device__ __forceinline__ uint32_t diag_of(uint32_t q, int32_t so) const
    {
        int32_t d = (int32_t)q - so % D;
        return (uint32_t)(d < 0 ? d + D : d);
    }

    // Two-hit decisions for one set of hits (at most one per lane; active
    // lanes are in subject order by lane).  Hits of the set that share a
    // diagonal are taken one per pass, lowest lane first (LDS atomic max on
    // the marker), so every diagonal sees its hits in subject order.  A hit
    // whose diagonal waits on a pending extension joins the blocked list.
    // Per hit (state per diagonal, flagged = extended to the right):
    //   pending            -> wait
    //   flagged            -> s < last: skip; else last = s, unflag
    //   diff >= window     -> last = s
    //   diff < 3           -> skip
    //   otherwise          -> extension (diagonal pending until it ran)
    __device__ __forceinline__ void decide(bool active, uint32_t q, int32_t so, uint32_t d)
    {
        for (;;) {
            if (!__ballot(active)) break;
            if (++gen == (1u << 26)) {
                for (int i = lane; i < D; i += HB_WAVE) marker[i] = 0;
                wave_lds_sync();
                gen = 1;
            }
            const uint32_t key = (gen << 6) | (uint32_t)(63 - lane);
            if (active) atomicMax(&marker[d], key);
            wave_lds_sync();
            const uint32_t mk = marker[d];
            const int32_t st = state[d];
            const bool win = active && mk == key;
            const uint32_t ust = (uint32_t)st;
            const bool pend = (ust & kPend) != 0;
            const bool flag = st < 0;
            const int32_t sb = so + base;
            const int32_t diff = (int32_t)((uint32_t)sb - ust);  // s - last if unflagged
            const bool job = win && !pend && !flag && diff >= 3 && diff < window;
            const bool wait = win && pend;
            int32_t nst;
            if (flag) nst = sb >= (int32_t)(ust & kVal) ? sb : st;
            else nst = diff >= window ? sb : (job ? (int32_t)kPend : st);
            if (win && !pend && nst != st) state[d] = nst;
            const uint64_t mj = __ballot(job), mb = __ballot(wait);
            if (job) {
                const uint32_t delta = (uint32_t)diff;  // 3..window-1
                jobs[n_jobs + lane_rank(mj)] = make_uint2(q | (delta << 24), (uint32_t)so);
            }
            if (wait) blk[n_blk + lane_rank(mb)] = make_uint2(q, (uint32_t)so);
            n_jobs += (uint32_t)__popcll(mj);
            n_blk += (uint32_t)__popcll(mb);
            n_ext += job ? 1u : 0u;
            active = active && !win;
            wave_lds_sync();
        }
    }

    // Runs every queued extension (one per lane), writes HSPs with score >=
    // cutoff and the diagonals' new states.
    template <typename SP>
    __device__ __forceinline__ void run_jobs(SP S, int32_t slen, uint32_t oid)
    {
        for (uint32_t j0 = 0; j0 < n_jobs; j0 += HB_WAVE) {
            const uint32_t j = j0 + (uint32_t)lane;
            const bool v = j < n_jobs;
            ext_result r = {0, 0, 0, 0, 0, 0};
            int32_t k = 0, q = 0, so = 0;
            if (v) {
                const uint2 jb = jobs[j];
                q = (int32_t)(jb.x & 0xffffffu);
                so = (int32_t)jb.y;
                const int32_t last_s = so - (int32_t)(jb.x >> 24);
                k = context_of(ctx, a->n_ctx, q);
                r = ext_two_hit(mat, S, slen, qry, a->qlen, last_s + 3, so, q, ctx[a->ctx_pad + k]);
                state[diag_of((uint32_t)q, so)] =
                    r.right ? (int32_t)(kFlag | (uint32_t)(r.s_last - 2 + base)) : so + base;
            }
            const bool emit = v && r.score >= ctx[2 * a->ctx_pad + k];
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
                        o.context = k;
                        o.q_start = r.q_start;
                        o.s_start = r.s_start;
                        o.q_off = q;
                        o.s_off = so;
                        o.len = r.len;
                        o.score = r.score;
                        a->out[slot] = o;
                    }
                }
            }
            n_hsp += emit ? 1u : 0u;
        }
        n_jobs = 0;
        wave_lds_sync();
    }

    // Re-decides the blocked hits in their original order (in place: each
    // round reads its 64 entries before any new entry is written, and new
    // entries never outnumber the ones read).
    __device__ __forceinline__ void replay()
    {
        const uint32_t nb = n_blk;
        n_blk = 0;
        for (uint32_t b0 = 0; b0 < nb; b0 += HB_WAVE) {
            const uint32_t b = b0 + (uint32_t)lane;
            const bool v = b < nb;
            const uint2 e = v ? blk[b] : make_uint2(0, 0);
            wave_lds_sync();
            decide(v, e.x, (int32_t)e.y, diag_of(e.x, (int32_t)e.y));
        }
    }

    template <typename SP>
    __device__ __forceinline__ void drain(SP S, int32_t slen, uint32_t oid)
    {
        while (n_jobs) {
            run_jobs(S, slen, oid);
            replay();
        }
    }

    template <typename SP>
    __device__ __forceinline__ void subject(SP S, int32_t slen, uint32_t oid)
    {
        uint32_t hits = 0;
        n_ext = 0;
        n_hsp = 0;
        const uint32_t n_entries = a->n_entries;
        int32_t s0m = 0;  // s0 mod D
        for (int32_t s0 = 0; s0 + 3 <= slen; s0 += HB_WAVE) {
            const int32_t s = s0 + lane;
            uint32_t cnt = 0, cst = 0;
            if (s + 3 <= slen) {
                const uint32_t idx = ((uint32_t)S[s] << 10) | ((uint32_t)S[s + 1] << 5) | S[s + 2];
                if ((pv[idx >> 5] >> (idx & 31)) & 1u) {
                    const uint2 c = a->cells[idx];
                    cst = c.x;
                    cnt = c.y;
                }
            }
            const uint32_t incl = wave_incl_add(cnt);
            const uint32_t excl = incl - cnt;
            const uint32_t H = lane63(incl);
            const int32_t ebase = (int32_t)cst - (int32_t)excl;  // entry index = flat index + ebase
            hits += H;
            uint32_t carry = 0;
            for (uint32_t r0 = 0; r0 < H; r0 += HB_WAVE) {
                if (n_jobs >= HB_WAVE || n_blk > (uint32_t)(kBlkCap - HB_WAVE)) drain(S, slen, oid);
                if (++rgen == (1u << 25)) {
                    seg[lane] = 0;
                    rgen = 1;
                }
                const uint32_t tag = rgen << 7;
                if (cnt && excl >= r0 && excl < r0 + HB_WAVE) seg[excl - r0] = tag | ((uint32_t)lane + 1);
                wave_lds_sync();
                const uint32_t sv = seg[lane];
                const uint32_t v = wave_incl_max((sv & ~127u) == tag ? (sv & 127u) : 0u);
                const uint32_t p = min(v ? v - 1 : carry, 63u);
                carry = lane63(p);
                const uint32_t h = r0 + (uint32_t)lane;
                const bool active = h < H;
                const int32_t eb = __shfl(ebase, (int)p, HB_WAVE);
                const uint32_t ei = min((uint32_t)((int32_t)h + eb), n_entries - 1);
                const uint32_t q = active ? a->entries[ei] : 0u;
                int32_t smp = s0m + (int32_t)p;
                if (smp >= D) smp -= D;
                int32_t d = (int32_t)q - smp;
                if (d < 0) d += D;
                wave_lds_sync();
                decide(active, q, s0 + (int32_t)p, (uint32_t)d);
            }
            s0m += HB_WAVE;
            if (s0m >= D) s0m -= D;
        }
        drain(S, slen, oid);
        const uint32_t e = wave_sum(n_ext), hs = wave_sum(n_hsp);
        if (lane == 0) {
            a->stats[3 * (size_t)oid + 0] = hits;
            a->stats[3 * (size_t)oid + 1] = e;
            a->stats[3 * (size_t)oid + 2] = hs;
        }
        if (base + slen + window >= (1 << 29)) {
            for (int i = lane; i < D; i += HB_WAVE) state[i] = 0;
            wave_lds_sync();
            base = window;
        } else {
            base += slen + window;
        }
    }
};

__global__ __launch_bounds__(HB_BLOCK) void hb_scan_kernel(hb_scan_args a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    const int lane = threadIdx.x & (HB_WAVE - 1);
    const int wid = threadIdx.x / HB_WAVE;
    int8_t *sm_mat = reinterpret_cast<int8_t *>(lds);
    uint32_t *sm_pv = reinterpret_cast<uint32_t *>(lds + 1024);
    int32_t *sm_ctx = reinterpret_cast<int32_t *>(lds + 1024 + 4 * HB_PV_WORDS);
    uint8_t *sm_q = reinterpret_cast<uint8_t *>(sm_ctx + 3 * a.ctx_pad);
    unsigned char *wb = sm_q + a.qlen_pad + (size_t)wid * a.wave_lds;
    const int32_t D = a.diag_len;

    for (int i = threadIdx.x; i < 256; i += blockDim.x)
        reinterpret_cast<uint32_t *>(sm_mat)[i] = reinterpret_cast<const uint32_t *>(a.mat32)[i];
    for (int i = threadIdx.x; i < HB_PV_WORDS; i += blockDim.x) sm_pv[i] = a.pv[i];
    for (int i = threadIdx.x; i < a.qlen_pad / 4; i += blockDim.x)
        reinterpret_cast<uint32_t *>(sm_q)[i] = reinterpret_cast<const uint32_t *>(a.query)[i];
    for (int i = threadIdx.x; i < a.n_ctx; i += blockDim.x) {
        sm_ctx[i] = a.ctx_off[i];
        sm_ctx[a.ctx_pad + i] = a.ctx_xdrop[i];
        sm_ctx[2 * a.ctx_pad + i] = a.ctx_cutoff[i];
    }

    WaveScan w;
    w.a = &a;
    w.mat = sm_mat;
    w.qry = sm_q;
    w.ctx = sm_ctx;
    w.pv = sm_pv;
    w.state = reinterpret_cast<int32_t *>(wb);
    w.marker = reinterpret_cast<uint32_t *>(wb + 4 * (size_t)D);
    w.seg = reinterpret_cast<uint32_t *>(wb + 8 * (size_t)D);
    w.jobs = reinterpret_cast<uint2 *>(wb + 8 * (size_t)D + 256);
    w.blk = w.jobs + kJobCap;
    w.sbuf = reinterpret_cast<uint8_t *>(w.blk + kBlkCap);
    w.lane = lane;
    w.D = D;
    w.window = a.window;
    w.base = a.window;
    w.gen = 0;
    w.rgen = 0;
    w.n_jobs = 0;
    w.n_blk = 0;
    for (int i = lane; i < 2 * D + 64; i += HB_WAVE) w.state[i] = 0;   // state, marker, seg
    __syncthreads();

    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(a.tile_counter, 1u);
        t = (uint32_t)__builtin_amdgcn_readfirstlane((int)t);
        if (t >= a.n_tiles) break;
        const uint32_t i_end = a.tiles[t + 1];
        for (uint32_t it = a.tiles[t]; it < i_end; ++it) {
            const uint32_t oid = a.order[it];
            const uint64_t sb = a.offs[oid];
            const int32_t slen = (int32_t)(a.offs[oid + 1] - sb);
            const uint8_t *S = a.res + sb;
            if (slen <= kSCap) {
                for (int i = lane; i < slen; i += HB_WAVE) w.sbuf[i] = S[i];
                wave_lds_sync();
                w.subject(w.sbuf, slen, oid);
                wave_lds_sync();
            } else {
                w.subject(S, slen, oid);
            }
        }
    }
}

// Plain 16-byte-per-lane copy: the known-size reference kernel for
// calibrating HBM byte counters and the achievable copy bandwidth.
__global__ __launch_bounds__(256) void copy_kernel(uint4 *dst, const uint4 *src, uint64_t n16)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
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

hipError_t hb_launch_lut_count(const uint8_t *qflag, int32_t qlen, const int32_t *,
                               const int8_t *mat32, int32_t threshold, uint32_t *count,
                               hipStream_t st)
{
    const size_t lds = 1024 + (size_t)((qlen + 15) & ~15);
    hipLaunchKernelGGL(lut_kernel<false>, dim3(HB_NCELLS / 256), dim3(256), lds, st, qflag, qlen,
                       mat32, threshold, count, nullptr, nullptr, nullptr, nullptr);
    return hipGetLastError();
}

hipError_t hb_launch_lut_scan(uint32_t *count_to_start, uint32_t *total, hipStream_t st)
{
    hipLaunchKernelGGL(lut_scan_kernel, dim3(1), dim3(1024), 0, st, count_to_start, total);
    return hipGetLastError();
}

hipError_t hb_launch_lut_fill(const uint8_t *qflag, int32_t qlen, const int32_t *,
                              const int8_t *mat32, int32_t threshold, const uint32_t *start,
                              uint32_t *entries, uint2 *cells, uint32_t *pv, hipStream_t st)
{
    const size_t lds = 1024 + (size_t)((qlen + 15) & ~15);
    hipLaunchKernelGGL(lut_kernel<true>, dim3(HB_NCELLS / 256), dim3(256), lds, st, qflag, qlen,
                       mat32, threshold, nullptr, start, entries, cells, pv);
    return hipGetLastError();
}

hipError_t hb_launch_scan(const hb_scan_args &a, int n_blocks, int block_threads, size_t lds_bytes,
                          hipStream_t st)
{
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(hb_scan_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    if (block_threads < HB_WAVE || block_threads > HB_BLOCK || block_threads % HB_WAVE)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(hb_scan_kernel, dim3(n_blocks), dim3(block_threads), lds_bytes, st, a);
    return hipGetLastError();
}

hipError_t hb_scan_occupancy(size_t lds_bytes, int block_threads, int *blocks_per_cu)
{
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(hb_scan_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, hb_scan_kernel,
                                                        block_threads, lds_bytes);
}
