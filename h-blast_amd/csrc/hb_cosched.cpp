// hb_cosched.cpp -- the co-scheduled volume loop (host only, no HIP): each
// volume of an H-BLAST database is split between the device (a prefix of
// its .hDb) and the caller's CPU search (the rest), and the device works on
// volume v + 1 while the CPU searches volume v, in the order of
// GPU_BLAST_PreliminarySearchEngine (ncbi_blast_files/functions.cpp:
// 1409-1790): fetch volume v's device results, update the ratio from the
// last device and CPU times (get_new_gpudb_ratio, :51-140), cut and launch
// volume v + 1 (get_GPU_DB_information, :449-571, after the memory check
// that sends a share that does not fit to the CPU whole, :1518-1530,
// 1584-1585), then run the CPU search of volume v's rest.  The device steps
// come in as callbacks (hb_cosearch binds the library's own).
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "hb_seed.h"

namespace {

// Reads up to `bytes` bytes of a file (all of it with bytes = UINT64_MAX).
int read_file(const std::string &path, uint64_t bytes, std::vector<uint8_t> &out)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return HB_ERR_ARG;
    if (bytes == UINT64_MAX) {
        if (std::fseek(f, 0, SEEK_END) != 0) {
            std::fclose(f);
            return HB_ERR_ARG;
        }
        const long end = std::ftell(f);
        std::rewind(f);
        if (end < 0) {
            std::fclose(f);
            return HB_ERR_ARG;
        }
        bytes = (uint64_t)end;
    }
    out.resize(bytes);
    const size_t got = bytes ? std::fread(out.data(), 1, bytes, f) : 0;
    std::fclose(f);
    return got == bytes ? HB_OK : HB_ERR_FORMAT;
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" int hb_cosched_run(const char *index_path, const hb_runtime_options *options,
                              int32_t cpu_threads, int32_t gapped, int32_t query_length,
                              uint32_t avg_query_length, const hb_cosched_steps *dev,
                              hb_cpu_search_fn cpu, void *cpu_user, hb_volume_report *reports,
                              int32_t report_cap, int32_t *n_volumes)
{
    if (!index_path || !dev || !dev->fits || !dev->launch || !dev->fetch || !cpu ||
        !n_volumes || cpu_threads < 1 || report_cap < 0 || (report_cap > 0 && !reports))
        return HB_ERR_ARG;
    *n_volumes = 0;
    std::vector<uint8_t> text;
    int rc = read_file(index_path, UINT64_MAX, text);
    if (rc) return rc;
    hb_hdb_index idx;
    uint64_t n_gl = 0, err_off = 0;
    rc = hb_hdbi_parse(reinterpret_cast<const char *>(text.data()), text.size(), &idx, nullptr, 0,
                       nullptr, 0, &n_gl, &err_off);
    if (rc) return rc;
    std::vector<hb_hdb_volume> vols((size_t)idx.n_volumes);
    std::vector<int32_t> gl((size_t)(n_gl ? n_gl : 1));
    rc = hb_hdbi_parse(reinterpret_cast<const char *>(text.data()), text.size(), &idx, vols.data(),
                       idx.n_volumes, gl.data(), gl.size(), &n_gl, &err_off);
    if (rc) return rc;
    const int32_t nv = idx.n_volumes;
    *n_volumes = nv;
    if (report_cap < nv) return HB_ERR_CAPACITY;
    std::string dir(index_path);
    const size_t slash = dir.find_last_of('/');
    dir = slash == std::string::npos ? std::string() : dir.substr(0, slash + 1);

    hb_runtime_options opts;
    if (options) opts = *options;
    else hb_runtime_options_default(&opts);
    float ratio = opts.gpu_ratio;
    int32_t fixed = opts.fixed_ratio;
    uint64_t ring_bytes[3] = {0, 0, 0};
    float ring_ratio[2] = {opts.gpu_ratio, 0.f};
    int64_t ring_hsp[2] = {0, 0};
    uint64_t seq = 0;          // order of launches and CPU searches
    bool running = false;      // a share is on the device
    std::vector<uint8_t> buf;
    for (int32_t v = 0; v < nv; ++v) std::memset(&reports[v], 0, sizeof(hb_volume_report));
    auto fail = [&](int r) {
        if (running && dev->abandon) dev->abandon(dev->user);
        return r;
    };
    for (int32_t v = -1; v < nv; ++v) {
        if (v >= 0) {   // volume v's device results
            hb_volume_report &r = reports[v];
            if (running) {
                running = false;
                rc = dev->fetch(dev->user, v, r.oid_base, &r.gpu_seconds, &r.n_hsp);
                if (rc) return fail(rc);
            }
            ring_hsp[v & 1] = (int64_t)r.n_hsp;
        }
        if (v + 1 < nv) {   // cut and launch volume v + 1
            const int32_t w = v + 1;
            const hb_hdb_volume &vol = vols[(size_t)w];
            ring_bytes[w % 3] = vol.bytes + vol.cpu_rest_bytes;
            // once a volume's ratio is 0 (the device got no share) the updates
            // stop and the CPU takes every later volume: the reference's update
            // divides by that ratio and stops the program there
            // (functions.cpp:65-69)
            const bool no_share = ring_ratio[0] == 0.f || ring_ratio[1] == 0.f;
            if (v > 0 && fixed != 1 && !no_share) {
                hb_balance_in in;
                std::memset(&in, 0, sizeof(in));
                in.volume = (uint32_t)v;
                in.ratio = ratio;
                in.cpu_threads = (uint32_t)cpu_threads;
                in.gpu_cards = 1;
                for (int k = 0; k < 3; ++k) in.volume_bytes[k] = ring_bytes[k];
                in.ratio_ring[0] = ring_ratio[0];
                in.ratio_ring[1] = ring_ratio[1];
                in.hsp_ring[0] = ring_hsp[0];
                in.hsp_ring[1] = ring_hsp[1];
                in.gapped = gapped ? 1 : 0;
                in.query_length = query_length;
                in.avg_query_length = avg_query_length;
                in.gpu_seconds = reports[v].gpu_seconds;
                in.cpu_seconds = reports[v - 1].cpu_seconds;
                rc = hb_balance_next_ratio(&opts, &in, &ratio);
                if (rc) return fail(rc);
            }
            hb_hdb_share share;
            rc = hb_hdb_split(vol.group_len, vol.n_groups, vol.gpu_seqs, vol.bytes, idx.blocksx,
                              idx.threadsx, cpu_threads, &ratio, &fixed, &share);
            if (rc) return fail(rc);
            ring_ratio[w & 1] = ratio;
            hb_volume_report &r = reports[w];
            r.oid_base = (int64_t)vol.break_limit - vol.volume_seqs;
            r.volume_seqs = vol.volume_seqs;
            r.gpu_seqs = share.gpu_seqs;
            r.ratio = ratio;
            if (share.gpu_seqs > 0) {
                int32_t fits = 1;
                const int32_t max_len = share.n_groups > 0 ? vol.group_len[share.n_groups - 1] : 0;
                rc = dev->fits(dev->user, w, &share, max_len, &r.need_bytes, &fits);
                if (rc) return fail(rc);
                if (!fits) {
                    r.gpu_seqs = 0;
                    r.on_cpu = 1;
                }
            }
            if (r.gpu_seqs > 0) {
                rc = read_file(dir + std::string(vol.name, vol.name_len), share.bytes, buf);
                if (rc) return fail(rc);
                r.launch_seq = ++seq;
                rc = dev->launch(dev->user, w, buf.data(), &share, idx.blocksx, idx.threadsx,
                                 vol.group_len);
                if (rc == HB_ERR_NOMEM) {   // the upload found no room: the CPU takes it
                    r.gpu_seqs = 0;
                    r.on_cpu = 1;
                } else if (rc) {
                    return fail(rc);
                } else {
                    running = true;
                }
            }
        }
        if (v >= 0) {   // the CPU search of volume v's rest
            hb_volume_report &r = reports[v];
            r.cpu_seq = ++seq;
            const auto t0 = std::chrono::steady_clock::now();
            rc = cpu(cpu_user, v, r.oid_base + r.gpu_seqs, r.oid_base + r.volume_seqs);
            r.cpu_seconds = seconds_since(t0);
            if (rc) return fail(HB_ERR_STATE);
        }
    }
    return HB_OK;
}
