// hb_seq.cpp -- host-only sequence helpers of the C ABI (no HIP): protein
// letters <-> ncbistdaa codes, protein FASTA parsing, six-frame BLASTX query
// construction.  Kept free of device code so that they also build with
// plain g++ (the sanitizer tests compile them with ASan/UBSan).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <vector>

#include "hb_seed.h"

extern "C" {

static const char kStdaa[] = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";

int hb_encode_protein(const char *text, int32_t len, uint8_t *out)
{
    if (!text || !out || len < 0) return HB_ERR_ARG;
    for (int32_t i = 0; i < len; ++i) {
        char ch = text[i];
        if (ch >= 'a' && ch <= 'z') ch = (char)(ch - 'a' + 'A');
        const char *p = strchr(kStdaa, ch);
        if (!p || ch == 0) return HB_ERR_RESIDUE;
        out[i] = (uint8_t)(p - kStdaa);
    }
    return HB_OK;
}

int hb_fasta_parse(const char *text, uint64_t len, uint8_t *residues, uint64_t res_cap,
                   uint64_t *offsets, uint64_t *titles, uint32_t *titles_len,
                   uint32_t seq_cap, uint64_t *n_res, uint32_t *n_seq)
{
    if (!n_res || !n_seq || (len && !text)) return HB_ERR_ARG;
    *n_res = 0;
    *n_seq = 0;
    // letter -> code + 1 (0: not a residue)
    uint8_t code[256] = {0};
    for (int i = 0; kStdaa[i]; ++i) {
        const unsigned char ch = (unsigned char)kStdaa[i];
        code[ch] = (uint8_t)(i + 1);
        if (ch >= 'A' && ch <= 'Z') code[ch - 'A' + 'a'] = (uint8_t)(i + 1);
    }
    const bool fill = residues || offsets;
    uint64_t nr = 0;
    uint64_t ns = 0;
    bool overflow = false;
    uint64_t i = 0;
    while (i < len) {
        uint64_t e = i;
        while (e < len && text[e] != '\n') ++e;   // line [i, e)
        const char c0 = text[i];
        if (c0 == '>') {
            if (ns >= 0xffffffffull) return HB_ERR_CAPACITY;
            if (fill) {
                if (ns < seq_cap) {
                    if (offsets) offsets[ns] = nr;
                    uint64_t te = e;
                    if (te > i + 1 && text[te - 1] == '\r') --te;
                    if (titles) titles[ns] = i + 1;
                    if (titles_len) titles_len[ns] = (uint32_t)std::min<uint64_t>(te - i - 1, 0xffffffffu);
                } else {
                    overflow = true;
                }
            }
            ++ns;
        } else if (c0 != ';') {
            for (uint64_t k = i; k < e; ++k) {
                const unsigned char ch = (unsigned char)text[k];
                const uint8_t cd = code[ch];
                if (cd) {
                    if (ns == 0) {   // residues before the first '>' line
                        *n_res = k;
                        return HB_ERR_RESIDUE;
                    }
                    if (fill) {
                        if (residues && nr < res_cap) residues[nr] = (uint8_t)(cd - 1);
                        else overflow = true;
                    }
                    ++nr;
                } else if (!(ch == ' ' || ch == '\t' || ch == '\r' || (ch >= '0' && ch <= '9'))) {
                    *n_res = k;
                    return HB_ERR_RESIDUE;
                }
            }
        }
        i = e + 1;
    }
    *n_res = nr;
    *n_seq = (uint32_t)ns;
    if (fill && offsets) {
        if (ns < seq_cap + 1ull) offsets[ns] = nr;
        else overflow = true;
    }
    return overflow ? HB_ERR_CAPACITY : HB_OK;
}

// Standard genetic code (NCBI translation table 1), codons indexed
// 16*b1 + 4*b2 + b3 with T=0, C=1, A=2, G=3.
static const char kCode1[] =
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG";

static int nt_index(char ch)
{
    switch (ch) {
    case 'T': case 't': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'A': case 'a': return 2;
    case 'G': case 'g': return 3;
    default: return -1;
    }
}

int32_t hb_blastx_concat_len(int32_t nt_len)
{
    if (nt_len < 0) return 0;
    int32_t total = 0;
    for (int f = 0; f < 3; ++f) total += 2 * (nt_len > f ? (nt_len - f) / 3 : 0);
    return total + 5;
}

int hb_blastx_translate(const char *nt, int32_t nt_len, uint8_t *out, int32_t *ctx_off,
                        int32_t *ctx_len)
{
    if (!nt || nt_len < 0 || !out || !ctx_off || !ctx_len) return HB_ERR_ARG;
    // fwd: plus strand; rev: reverse complement (T<->A, C<->G)
    static const int comp[4] = {2, 3, 0, 1};
    std::vector<int> fwd((size_t)nt_len), rev((size_t)nt_len);
    for (int32_t i = 0; i < nt_len; ++i) {
        fwd[i] = nt_index(nt[i]);
        const int x = nt_index(nt[nt_len - 1 - i]);
        rev[i] = x < 0 ? -1 : comp[x];
    }
    int32_t pos = 0;
    for (int fr = 0; fr < 6; ++fr) {
        const std::vector<int> &src = fr < 3 ? fwd : rev;
        const int f = fr % 3;
        const int32_t n = nt_len > f ? (nt_len - f) / 3 : 0;
        if (fr) out[pos++] = 0;  // sentinel between contexts
        ctx_off[fr] = pos;
        ctx_len[fr] = n;
        for (int32_t i = 0; i < n; ++i) {
            const int b1 = src[f + 3 * i], b2 = src[f + 3 * i + 1], b3 = src[f + 3 * i + 2];
            uint8_t aa = 21;  // X
            if (b1 >= 0 && b2 >= 0 && b3 >= 0) {
                const char ch = kCode1[16 * b1 + 4 * b2 + b3];
                aa = (uint8_t)(strchr(kStdaa, ch) - kStdaa);
            }
            out[pos++] = aa;
        }
    }
    return HB_OK;
}

}  // extern "C"
