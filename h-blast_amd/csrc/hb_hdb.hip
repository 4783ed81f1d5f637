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
// pass transposes 64 rows x 64 sequences through LDS so that every store
// instruction writes 64 consecutive bytes of one packed sequence.
#include <algorithm>

#include "hb_internal.h"

namespace {

constexpr int kThreads = 256;
constexpr uint32_t kRows = 64;     // rows per LDS tile
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
// the group length backwards.
__global__ void __launch_bounds__(kThreads) hdb_lengths_kernel(hb_hdb_args a)
{
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t c = blockIdx.x * kThreads + threadIdx.x; c < a.n_seqs; c += stride) {
        const Seat s = seat_of(a, c);
        const uint32_t t = c % a.threadsx;
        const uint32_t *w = reinterpret_cast<const uint32_t *>(a.hdb + s.block_off) + t;
        uint32_t len = 0;
        for (int32_t i = (int32_t)(s.L / 4) - 1; i >= 0; --i) {
            const uint32_t v = w[(uint64_t)i * a.threadsx];
            if (v) {
                len = 4u * (uint32_t)i + (32u - (uint32_t)__clz((int)v) + 7u) / 8u;
                break;
            }
        }
        a.lens[c] = len;
    }
}

// One work item = up to 64 consecutive sequences of one block.
__global__ void __launch_bounds__(kThreads) hdb_copy_kernel(hb_hdb_args a, uint32_t per_block,
                                                            uint32_t n_items)
{
    __shared__ uint32_t tile[kRows * kPitch];
    __shared__ uint32_t s_len[kCols];
    __shared__ uint64_t s_off[kCols];
    __shared__ uint32_t s_rows;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
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
            if (lane < nt)
                for (uint32_t r = wave; r < kRows && r0 + r < rows; r += kThreads / 64)
                    tile[r * kPitch + lane] = w[(uint64_t)(r0 + r) * a.threadsx + lane];
            __syncthreads();
            const uint32_t p0 = 4u * r0;
            for (uint32_t j = wave; j < nt; j += kThreads / 64) {
                const uint32_t len = s_len[j];
                if (len <= p0) continue;
                const uint32_t pend = min(len, p0 + 4u * kRows);
                uint8_t *dst = a.res + s_off[j];
                for (uint32_t p = p0 + lane; p < pend; p += 64)
                    dst[p] = (uint8_t)(tile[(p / 4u - r0) * kPitch + j] >> (8u * (p & 3u)));
            }
            __syncthreads();
        }
        __syncthreads();   // s_rows is read by every thread before the next item resets it
    }
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
