/*
 * hbo_hdb.c -- CPU oracle for H-BLAST's GPU database volume (.hDb) and its
 * text index (.hDbI) (test infrastructure only; see hbo.h).
 *
 * Writer: CreateH_BLAST_database (ncbi_blast_files/functions.cpp:785-1165)
 *   GPU share of a volume      functions.cpp:842-843  (99.9%, multiple of 32)
 *   group lengths              functions.cpp:844-903  (last member of each
 *                              group of blocksx*threadsx, Round2Multiple,
 *                              H-BLAST/h_blast.c:11-17, WIDTH_MULTIPLE 16,
 *                              H-BLAST/h_blast.h:4)
 *   file size                  functions.cpp:906-911
 *   byte placement             functions.cpp:1026-1073
 *   index text                 functions.cpp:257-309
 * Reader: the GPU kernel's subject walk (H-BLAST/h_blast_kernel.cu:593-608,
 * 1060-1064; the device copy of the file starts after threadsx ints of
 * padding, H-BLAST/h_blast.cu:993,1094), which the writer must agree with.
 *
 * Two cases the reference leaves undefined are fixed here (and in the
 * product): an empty sequence writes no byte (the reference stores
 * sequence[0] of it, functions.cpp:1058-1066), and a group-length vector
 * holds exactly one entry per group (the reference also stores one past
 * its end when the GPU share is a multiple of the group size,
 * functions.cpp:892-903).  The file does not record sequence lengths: the
 * reader takes a sequence to end after its last non-zero byte, so trailing
 * gap codes (0) are not recovered.
 */
#include "hbo.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

int32_t hbo_hdb_gpu_seqs(int32_t volume_seqs)
{
    /* functions.cpp:842-843, in the same double arithmetic */
    int32_t n = (int32_t)((99.9 * volume_seqs) / 100);
    return (n / 32) * 32;
}

static int32_t round16(int32_t x) { return x % 16 ? x + 16 - x % 16 : x; }

int hbo_hdb_layout(const uint64_t *offsets, int32_t n_seqs, int32_t blocksx, int32_t threadsx,
                   int32_t *group_len, uint64_t *bytes)
{
    if (n_seqs < 0 || blocksx < 1 || threadsx < 1) return HBO_ERR_ARG;
    const int64_t G = (int64_t)blocksx * threadsx;
    const int64_t n_groups = (n_seqs + G - 1) / G;
    uint64_t total = 0;
    for (int64_t g = 0; g < n_groups; ++g) {
        /* the group's last member is its longest (the volume is sorted) */
        int64_t last = (g + 1) * G - 1;
        if (last >= n_seqs) last = n_seqs - 1;
        const int32_t L = round16((int32_t)(offsets[last + 1] - offsets[last]));
        group_len[g] = L;
        const int64_t members = (g + 1) * G <= n_seqs ? G : n_seqs - g * G;
        /* a partial group occupies whole blocks of threadsx sequences */
        const int64_t slots = (members + threadsx - 1) / threadsx * threadsx;
        total += (uint64_t)slots * (uint64_t)L;
    }
    *bytes = total;
    return 0;
}

int hbo_hdb_pack(const uint8_t *residues, const uint64_t *offsets, int32_t n_seqs,
                 int32_t blocksx, int32_t threadsx, const int32_t *group_len, uint8_t *out,
                 uint64_t out_bytes)
{
    if (n_seqs < 0 || blocksx < 1 || threadsx < 1) return HBO_ERR_ARG;
    memset(out, 0, out_bytes);
    const int64_t G = (int64_t)blocksx * threadsx;
    uint64_t base = 0;
    for (int64_t c = 0; c < n_seqs; ++c) {
        const int64_t g = c / G, in_group = c % G;
        if (c > 0 && in_group == 0) base += (uint64_t)group_len[g - 1] * (uint64_t)G;
        const int64_t block = in_group / threadsx, t = in_group % threadsx;
        const uint64_t block_base = (uint64_t)block * threadsx * (uint64_t)group_len[g];
        const uint8_t *s = residues + offsets[c];
        const int64_t len = (int64_t)(offsets[c + 1] - offsets[c]);
        for (int64_t a = 0; a < len; ++a) {
            const uint64_t at = base + block_base + ((uint64_t)t + (uint64_t)(a / 4) * threadsx) * 4 +
                                (uint64_t)(a % 4);
            if (at >= out_bytes) return HBO_ERR_ARG;
            out[at] = s[a];
        }
    }
    return 0;
}

int hbo_hdb_unpack(const uint8_t *hdb, uint64_t bytes, int32_t n_seqs, int32_t blocksx,
                   int32_t threadsx, const int32_t *group_len, uint8_t *residues,
                   uint64_t *offsets)
{
    if (n_seqs < 0 || blocksx < 1 || threadsx < 1) return HBO_ERR_ARG;
    /* the kernel's view: 4-byte words; a thread starts at word tx of its
     * block (block_base = bx * threadsx * L / 4 words) and steps threadsx
     * words per 4 residues; after a group every pointer moves by
     * (group size - bx * threadsx) * L / 4 words, i.e. to the same block of
     * the next group */
    const int64_t G = (int64_t)blocksx * threadsx;
    const uint32_t *w = (const uint32_t *)hdb;
    uint64_t group_word = 0, pos = 0;
    offsets[0] = 0;
    for (int64_t c = 0; c < n_seqs; ++c) {
        const int64_t g = c / G, in_group = c % G;
        if (c > 0 && in_group == 0) group_word += (uint64_t)group_len[g - 1] / 4 * (uint64_t)G;
        const int64_t bx = in_group / threadsx, tx = in_group % threadsx;
        const int32_t L = group_len[g];
        const uint64_t first = group_word + (uint64_t)bx * threadsx * (uint64_t)(L / 4) + tx;
        if (L > 0 && (first + (uint64_t)(L / 4 - 1) * threadsx + 1) * 4 > bytes) return HBO_ERR_ARG;
        int32_t len = 0;
        for (int32_t i = 0; i < L / 4; ++i) {
            const uint32_t v = w[first + (uint64_t)i * threadsx];
            for (int b = 0; b < 4; ++b)
                if ((v >> (8 * b)) & 0xffu) len = 4 * i + b + 1;
        }
        if (residues)
            for (int32_t a = 0; a < len; ++a)
                residues[pos + (uint64_t)a] =
                    (uint8_t)(w[first + (uint64_t)(a / 4) * threadsx] >> (8 * (a % 4)));
        pos += (uint64_t)len;
        offsets[c + 1] = pos;
    }
    return 0;
}

int64_t hbo_hdbi_format(int32_t blocksx, int32_t threadsx, int32_t percentage, int32_t n_volumes,
                        const hbo_hdb_volume *vols, char *text, uint64_t cap)
{
    /* functions.cpp:257-272 then :276-309 per volume */
    uint64_t used = 0;
    char line[512];
#define HBO_PUT(...)                                                       \
    do {                                                                   \
        int k_ = snprintf(line, sizeof line, __VA_ARGS__);                 \
        if (k_ < 0) return HBO_ERR_ARG;                                    \
        if (text && used + (uint64_t)k_ <= cap) memcpy(text + used, line, (size_t)k_); \
        used += (uint64_t)k_;                                              \
    } while (0)
    HBO_PUT("#Number of blocksx\n%d\n", blocksx);
    HBO_PUT("#Number of threadsx\n%d\n", threadsx);
    HBO_PUT("#Percentage\n%d\n", percentage);
    HBO_PUT("#Number of volumes\n%d\n", n_volumes);
    for (int32_t v = 0; v < n_volumes; ++v) {
        const hbo_hdb_volume *V = &vols[v];
        HBO_PUT("#Volume name\n%s\n", V->name);
        HBO_PUT("#Volume size in bytes\n%llu\n", (unsigned long long)V->bytes);
        HBO_PUT("#Volume sequences\n%d\n", V->volume_seqs);
        HBO_PUT("#Number of sequences in GPU volume\n%d\n", V->gpu_seqs);
        HBO_PUT("#Store limit\n%d\n", V->store_limit);
        HBO_PUT("#Break limit\n%d\n", V->break_limit);
        HBO_PUT("#Group number\n%d\n", V->n_groups);
        HBO_PUT("#Length of each group\n");
        for (int32_t g = 0; g < V->n_groups; ++g) HBO_PUT("%d\n", V->group_len[g]);
        HBO_PUT("#Length of the rest sequences in CPU database\n%llu\n",
                (unsigned long long)V->cpu_rest_bytes);
    }
#undef HBO_PUT
    return (int64_t)used;
}
