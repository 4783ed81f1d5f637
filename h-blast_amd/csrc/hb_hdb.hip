// Unpacking of an H-BLAST .hDb volume into the packed database layout
// (residues back to back, one offset per sequence).  Pure byte movement,
// bound by HBM bandwidth: the lengths pass reads each sequence's padding
// tail once, the copy pass reads its residues and writes them packed.
//
// A block of threadsx sequences stores row i (residues 4i..4i+3 of every
// sequence) as threadsx consecutive 4-byte words (the stripes of
// ncbi_blast_files/functions.cpp:1058-1066, read the same way by
// H-BLAST/h_blast_kernel.cu:593-608), so consecutive lanes reading
// consecutive sequences of a row make whole-row, coalesced loads.  The copy
// pass transposes 64 rows x 64 sequences through LDS so that a store
// instruction writes 256 consecutive, 4-aligned bytes of one packed
// sequence.
#include <algorithm>

#include "hb_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kRows = 64;     // rows per LDS tile (one more is staged)
constexpr uint32_t kCols = 64;     // sequences per work item
constexpr uint32_t kPitch = 65;    // LDS words per tile row (bank spread)

struct Seat {
    uint32_t g;          // group
    uint64_t block_off;  // byte offset of the sequence's block
    uint32_t L;          // group length
};

__device__ __forceinline__ Seat seat_of(const hb_hdb_args &a, uint32_t c)
{
    const uint32_t B = c / a.threadsx;
    const uint32_t g = B / a.blocksx;
    const uint32_t b = B - g * a.blocksx;
    const uint32_t L = (uint32_t)a.group_len[g];
    return {g, a.group_base[g] + (uint64_t)b * a.threadsx * L, L};
}

// One thread per sequence: the last non-zero byte, searched from the end of
// the group length backwards, eight stripes (independent loads) at a time.
__global__ void __launch_bounds__(kThreads) hdb_lengths_kernel(hb_hdb_args a)
{
    constexpr int kStep = 8;
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c < a.n_seqs; c += stride) {
        const Seat s = seat_of(a, c);
        const uint32_t t = c % a.threadsx;
        const uint32_t *w = reinterpret_cast<const uint32_t *>(a.hdb + s.block_off) + t;
        uint32_t len = 0;
        for (int32_t top = (int32_t)(s.L / 4) - 1; top >= 0; top -= kStep) {
            uint32_t v[kStep];
#pragma unroll
            for (int k = 0; k < kStep; ++k)
                v[k] = top - k >= 0 ? w[(uint64_t)(top - k) * a.threadsx] : 0u;
            int hit = -1;
#pragma unroll
            for (int k = kStep - 1; k >= 0; --k)
                if (v[k]) hit = k;   // the highest stripe with a non-zero byte
            if (hit >= 0) {
                const uint32_t x = v[hit];
                len = 4u * (uint32_t)(top - hit) + (32u - (uint32_t)__clz((int)x) + 7u) / 8u;
                break;
            }
        }
        a.lens[c] = len;
    }
}

// One work item = up to 64 consecutive sequences of one block.  Per tile of
// 64 rows (plus the next row, for dwords that straddle two rows): every
// sequence's bytes whose packed address is 4-aligned go out as one dword
// per lane (a 256-byte store per sequence and tile, assembled from two
// stripes with alignbyte); its first (up to 3) bytes before the first
// aligned address and its last (up to 3) bytes after the last whole dword
// go out as single bytes, so no store touches a neighbouring sequence.
// Every copied residue is checked against the alphabet on the way.
// A wavefront issues all of its rows of a tile before writing them to LDS.
// (Also issuing the next tile's loads before the stores was measured
// slower: 0.614 vs 0.578 ms on the 2M-sequence volume, at 100 instead of 34
// VGPRs.)
__global__ void __launch_bounds__(kThreads) hdb_copy_kernel(hb_hdb_args a, uint32_t per_block,
                                                            uint32_t n_items)
{
    __shared__ uint32_t tile[(kRows + 1) * kPitch];
    __shared__ uint32_t s_len[kCols];
    __shared__ uint64_t s_off[kCols];
    __shared__ uint32_t s_rows;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    uint32_t bad = 0;   // a residue code >= HB_ALPHA was copied
    for (uint32_t k = blockIdx.x; k < n_items; k += gridDim.x) {
        const uint32_t B = k / per_block;
        const uint32_t t0 = (k - B * per_block) * kCols;
        const uint64_t c0 = (uint64_t)B * a.threadsx + t0;
        if (c0 >= a.n_seqs) continue;   // uniform over the workgroup
        const uint32_t nt = (uint32_t)min((uint64_t)min(kCols, a.threadsx - t0), a.n_seqs - c0);
        const Seat s = seat_of(a, (uint32_t)c0);
        const uint32_t *w = reinterpret_cast<const uint32_t *>(a.hdb + s.block_off) + t0;
        if (tid == 0) s_rows = 0;
        __syncthreads();
        if (tid < nt) {
            const uint32_t len = a.lens[c0 + tid];
            s_len[tid] = len;
            s_off[tid] = a.offs[c0 + tid];
            atomicMax(&s_rows, (len + 3u) / 4u);
        }
        __syncthreads();
        const uint32_t rows = s_rows;   // <= L / 4: every load below is inside the block
        for (uint32_t r0 = 0; r0 < rows; r0 += kRows) {
            if (lane < nt) {
                // all of the wave's rows of the tile in flight at once, then
                // into LDS (one load at a time left a single 256-byte load
                // per wavefront waiting on HBM latency)
                constexpr uint32_t kPer = (kRows + 1 + kThreads / 64 - 1) / (kThreads / 64);
                uint32_t v[kPer];
#pragma unroll
                for (uint32_t i = 0; i < kPer; ++i) {
                    const uint32_t r = wave + i * (kThreads / 64);
                    v[i] = r <= kRows && r0 + r < rows ? w[(uint64_t)(r0 + r) * a.threadsx + lane]
                                                       : 0u;
                }
#pragma unroll
                for (uint32_t i = 0; i < kPer; ++i) {
                    const uint32_t r = wave + i * (kThreads / 64);
                    if (r <= kRows && r0 + r < rows) tile[r * kPitch + lane] = v[i];
                }
            }
            __syncthreads();
            for (uint32_t j = wave; j < nt; j += kThreads / 64) {
                const uint32_t len = s_len[j];
                if (len <= 4u * r0) continue;
                const uint64_t off = s_off[j];
                const uint32_t h = (4u - (uint32_t)(off & 3u)) & 3u;   // head bytes
                // whole dwords m (bytes h + 4m .. h + 4m + 3, rows m and m + 1)
                const uint32_t M = len >= h ? (len - h) / 4u : 0u;
                const uint32_t m = r0 + lane;
                if (m < M) {
                    const uint32_t lo = tile[lane * kPitch + j];
                    const uint32_t v =
                        h ? __builtin_amdgcn_alignbyte(tile[(lane + 1) * kPitch + j], lo, h) : lo;
                    *reinterpret_cast<uint32_t *>(a.res + off + h + 4u * m) = v;
#pragma unroll
                    for (int b = 0; b < 4; ++b) bad |= ((v >> (8 * b)) & 0xffu) >= HB_ALPHA;
                }
                // lanes 0-3: head bytes (first tile); lanes 4-7: tail bytes
                // (the tile holding row M, where they start)
                const uint32_t head = min(h, len);
                const uint32_t tail0 = h + 4u * M;
                uint32_t p = 0xffffffffu;
                if (lane < 4u) {
                    if (r0 == 0 && lane < head) p = lane;
                } else if (lane < 8u && len >= h && M >= r0 && M < r0 + kRows) {
                    if (tail0 + (lane - 4u) < len) p = tail0 + (lane - 4u);
                }
                if (p != 0xffffffffu) {
                    const uint32_t r = (tile[(p / 4u - r0) * kPitch + j] >> (8u * (p & 3u))) & 0xffu;
                    a.res[off + p] = (uint8_t)r;
                    bad |= r >= HB_ALPHA;
                }
            }
            __syncthreads();
        }
        __syncthreads();   // s_rows is read by every thread before the next item resets it
    }
    const unsigned long long bm = __ballot(bad != 0);
    if (bm && lane == (uint32_t)__ffsll(bm) - 1u) atomicOr(a.bad, 1u);
}

}  // namespace

hipError_t hb_launch_hdb_lengths(const hb_hdb_args &a, hipStream_t st)
{
    if (a.n_seqs == 0) return hipSuccess;
    const uint32_t blocks = std::min((a.n_seqs + kThreads - 1) / kThreads, 65536u);
    hipLaunchKernelGGL(hdb_lengths_kernel, dim3(blocks), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t hb_launch_hdb_copy(const hb_hdb_args &a, hipStream_t st)
{
    if (a.n_seqs == 0) return hipSuccess;
    const uint32_t per_block = (a.threadsx + kCols - 1) / kCols;
    const uint64_t n_blocks = ((uint64_t)a.n_seqs + a.threadsx - 1) / a.threadsx;
    const uint64_t items = n_blocks * per_block;
    if (items > 0xffffffffull) return hipErrorInvalidValue;
    const uint32_t grid = (uint32_t)(items < (1u << 20) ? items : (1u << 20));
    hipLaunchKernelGGL(hdb_copy_kernel, dim3(grid), dim3(kThreads), 0, st, a, per_block,
                       (uint32_t)items);
    return hipGetLastError();
}
