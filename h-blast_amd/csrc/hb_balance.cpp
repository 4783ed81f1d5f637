// CPU/GPU split of a database volume (host code, no HIP).
//
// H-BLAST searches each volume on both sides at once: the device scans a
// prefix of the volume's .hDb while NCBI's CPU threads search the rest of
// the volume, and the size of that prefix moves from volume to volume so
// that both sides finish together (GPU_BLAST_PreliminarySearchEngine,
// ncbi_blast_files/functions.cpp:1409-1520).  Three host routines decide it:
//   - the runtime options file (Read_H_BLAST_runtime_options,
//     functions.cpp:142-203);
//   - the split of a volume at a workload ratio (get_GPU_DB_information,
//     functions.cpp:449-571): workload = group length x sequences; the GPU
//     keeps the .hDb's first whole groups and the first sequences of the
//     next group, so its share is a prefix of the file;
//   - the ratio for the next volume from the last wall times
//     (get_new_gpudb_ratio, functions.cpp:51-140).
// The arithmetic follows the reference's types step by step (float ratios,
// double workloads and times, int sequence counts), so the same inputs give
// the same ratios and the same split.  Built with -ffp-contract=off.
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "hb_seed.h"

extern "C" void hb_runtime_options_default(hb_runtime_options *o)
{
    if (!o) return;
    // the reference's values when no options file exists (functions.cpp:14-26,
    // H-BLAST/h_blast.cu:71)
    o->gpu_ratio = -1.0f;
    o->fixed_ratio = 1;
    o->capability_rate = 12.0;
    o->max_hits_per_subject = 19;
    o->mem_ecc = 0;
    o->load_balancer = 1;
    o->load_factor = 4;
    o->subject_chunk = 200;
    o->base_value = 100;
}

namespace {

// atof / atoi of one line (glibc: strtod, and (int) strtol base 10)
double line_atof(const std::string &s) { return std::strtod(s.c_str(), nullptr); }
int line_atoi(const std::string &s) { return (int)std::strtol(s.c_str(), nullptr, 10); }

}  // namespace

extern "C" int hb_runtime_options_parse(const char *text, uint64_t len, hb_runtime_options *o)
{
    if (!o || (len && !text)) return HB_ERR_ARG;
    // 18 lines: a comment line, then the value, for each of 9 options
    std::string lines[18];
    uint64_t pos = 0;
    for (int i = 0; i < 18; ++i) {
        if (pos >= len) return HB_ERR_FORMAT;
        const char *nl = (const char *)std::memchr(text + pos, '\n', len - pos);
        const uint64_t end = nl ? (uint64_t)(nl - text) : len;
        lines[i].assign(text + pos, end - pos);
        if (lines[i].find('\0') != std::string::npos) lines[i].resize(lines[i].find('\0'));
        pos = end + 1;
    }
    hb_runtime_options r;
    const double ratio = line_atof(lines[1]);
    r.gpu_ratio = (ratio < 0 || ratio > 1) ? 0.0f : (float)ratio;
    r.fixed_ratio = line_atoi(lines[3]) == 1 ? 1 : 0;
    const double cap = line_atof(lines[5]);
    r.capability_rate = cap <= 0 ? 12.0 : cap;
    const int hits = line_atoi(lines[7]);
    r.max_hits_per_subject = hits > 0 ? hits : 20;
    r.mem_ecc = line_atoi(lines[9]) != 0 ? 1 : 0;
    r.load_balancer = line_atoi(lines[11]) != 0 ? 1 : 0;
    const int lf = line_atoi(lines[13]);
    r.load_factor = lf < 0 ? 4 : lf;
    const int chunk = line_atoi(lines[15]);
    r.subject_chunk = chunk < 0 ? 200 : chunk;
    const int base = line_atoi(lines[17]);
    r.base_value = base < 0 ? 100 : base;
    *o = r;
    return HB_OK;
}

extern "C" int hb_hdb_split(const int32_t *group_len, int32_t n_groups, int32_t gpu_seqs,
                            uint64_t bytes, int32_t blocksx, int32_t threadsx,
                            int32_t cpu_threads, float *ratio, int32_t *fixed, hb_hdb_share *out)
{
    if (!ratio || !fixed || !out || blocksx < 1 || threadsx < 1 || gpu_seqs < 0 || n_groups < 0 ||
        (n_groups > 0 && !group_len))
        return HB_ERR_ARG;
    const int64_t G = (int64_t)blocksx * threadsx;
    if ((int64_t)n_groups != ((int64_t)gpu_seqs + G - 1) / G) return HB_ERR_ARG;
    if (!(*ratio == -1.0f || !(*ratio < 0))) return HB_ERR_ARG;   // -1, NaN or >= 0
    for (int32_t g = 0; g < n_groups; ++g)
        if (group_len[g] < 0) return HB_ERR_ARG;

    out->gpu_seqs = gpu_seqs;
    out->n_groups = n_groups;
    out->bytes = bytes;

    // ratio: the caller's, or a starting value by CPU thread count, which
    // also makes the ratio adaptive from 3 threads (functions.cpp:466-513)
    double r;
    if (*ratio == -1.0f) {
        switch (cpu_threads) {
        case 1: r = 1; break;
        case 2: r = 0.9; break;
        case 3: r = 0.9; *fixed = 0; break;
        case 4: r = 0.8; *fixed = 0; break;
        case 5: r = 0.75; *fixed = 0; break;
        case 6: r = 0.65; *fixed = 0; break;
        case 7: r = 0.55; *fixed = 0; break;
        case 8: r = 0.45; *fixed = 0; break;
        default: r = 0.4; *fixed = 0; break;
        }
    } else {
        r = (double)*ratio;
    }
    if (n_groups == 0) return HB_OK;   // an empty .hDb: nothing to split

    // workload = sum over groups of group length x sequences in the group
    const int32_t last = n_groups - 1;
    const double last_cost = (double)group_len[last] * (double)(gpu_seqs - (int64_t)last * G);
    double total = 0;
    for (int32_t l = 0; l < last; ++l) total += group_len[l];
    total *= (double)G;
    total += last_cost;
    const double target = total * r;

    // drop groups from the end while the kept ones weigh more than the target
    int32_t k = n_groups;
    double kept = total;
    if (r < 1) {
        kept -= last_cost;
        --k;
        while (k > 0 && kept > target) {
            kept -= (double)group_len[k - 1] * (double)G;
            --k;
        }
    }
    if (k == n_groups) return HB_OK;

    // then the first sequences of group k that fill the rest of the target:
    // (rest / (length x threadsx)) x threadsx, truncated to int, so not
    // always whole blocks (the bytes still cover whole blocks); rest is below
    // the dropped group's workload, so this stays below the group's size
    const double rest = target - kept;
    const int32_t Lk = group_len[k];
    const double fill = Lk != 0 ? rest / (double)((int64_t)Lk * threadsx) * (double)threadsx : 0;
    const int32_t added = fill < 2147483647.0 ? (int32_t)fill : 2147483647;
    kept += (double)Lk * (double)added;
    if (*fixed != 1) {
        float nr = (float)(kept / total);
        nr = nr > 1 ? 1 : nr;
        nr = nr < 0 ? 0 : nr;
        *ratio = nr;
    }
    if ((int64_t)k * G + added < (int64_t)gpu_seqs) {
        out->gpu_seqs = (int32_t)((int64_t)k * G + added);
        out->n_groups = k + (added != 0 ? 1 : 0);
        uint64_t b = 0;
        for (int32_t l = 0; l < k; ++l) b += (uint64_t)group_len[l] * (uint64_t)G;
        if (added)
            b += (uint64_t)Lk * (((uint64_t)added + threadsx - 1) / threadsx) * (uint64_t)threadsx;
        out->bytes = b;
    }
    return HB_OK;
}

extern "C" int hb_balance_next_ratio(const hb_runtime_options *o, const hb_balance_in *in,
                                     float *out)
{
    if (!o || !in || !out || in->volume < 1) return HB_ERR_ARG;
    const uint32_t v = in->volume;
    const uint64_t *B = in->volume_bytes;
    // the reference stops the program on these (functions.cpp:58-68)
    if (B[0] == 0 || B[1] == 0 || B[2] == 0) return HB_ERR_ARG;
    if (in->ratio_ring[0] == 0.0f || in->ratio_ring[1] == 0.0f) return HB_ERR_ARG;

    double gpu_t = in->gpu_seconds;
    double cpu_t = in->cpu_seconds;
    const uint32_t cur = v & 1u, prev = (v - 1) & 1u;
    double rate1 = (double)in->ratio_ring[cur] * 0.96;
    double rate2 = (double)in->ratio_ring[prev] * 0.96;

    // the CPU load of this volume and of the previous one: residues left to
    // the CPU, plus a charge per GPU HSP (the CPU re-extends them)
    if (in->gapped) {
        const uint32_t lpq = in->avg_query_length;
        const uint32_t m = 1500u > lpq ? lpq : 1500u;
        const double per_hsp =
            o->base_value + (double)(750 - o->base_value) * (double)m / 1500;
        rate1 = (double)B[v % 3] * (1.0 - rate1) + (double)in->hsp_ring[cur] * per_hsp / (rate1 + 0.03);
        rate2 = (double)B[(v - 1) % 3] * (1.0 - rate2) +
                (double)in->hsp_ring[prev] * per_hsp / (rate2 + 0.03);
    } else if (!o->load_balancer) {
        rate1 = (double)B[v % 3] * (1.0 - rate1);
        rate2 = (double)B[(v - 1) % 3] * (1.0 - rate2);
    } else {
        rate1 = (double)B[v % 3] * (1.0 - rate1) +
                (double)in->hsp_ring[cur] * (double)o->load_factor / (rate1 + 0.03);
        // the reference divides the previous volume's charge by the
        // updated rate1 here (functions.cpp:93); kept, it decides the ratio
        rate2 = (double)B[(v - 1) % 3] * (1.0 - rate2) +
                (double)in->hsp_ring[prev] * (double)o->load_factor / (rate1 + 0.03);
    }
    const double cpu_scale = rate1 / rate2;
    double gpu_scale = (double)B[(v + 1) % 3] / (double)B[v % 3];
    const int32_t ql = in->query_length;
    if (ql > 1000 && ql < 4000)
        gpu_scale = std::pow(gpu_scale, 0.9);
    else if (ql >= 4000 && ql < 8000)
        gpu_scale = std::pow(gpu_scale, 0.8);
    else if (ql >= 8000)
        gpu_scale = std::pow(gpu_scale, 0.75);
    cpu_t *= cpu_scale;
    gpu_t *= gpu_scale;

    if (gpu_t == 0) {
        *out = (float)0.9;
        return HB_OK;
    }
    const double cpu_threads = (double)in->cpu_threads, cards = (double)in->gpu_cards;
    const float equivalent = (float)(cpu_threads / o->capability_rate + cards);
    float adjust = (float)((cpu_t - gpu_t) / gpu_t);
    const double E = 0.65;
    double M = E * cards;
    M = M / (M + (1 - E) * cpu_threads / o->capability_rate);
    if (cpu_t - gpu_t > 0)
        adjust = (float)((double)adjust * (M / (double)equivalent));
    else
        adjust = (float)((double)adjust * (1.0 - M / (double)equivalent));
    float nr = (float)((double)in->ratio + (double)adjust / 0.96);
    nr = nr > 1 ? 1 : nr;
    nr = nr < 0 ? 0 : nr;
    *out = nr;
    return HB_OK;
}
