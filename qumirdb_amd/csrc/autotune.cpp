// The aggregate autotuner's cache and its C exports (autotune.h).
#include "autotune.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <mutex>
#include <tuple>

#include "../../include/qdb_gpu.h"

namespace qdbgpu {

namespace {

struct Entry {
    TuneConfig cfg;
    double ms;
};

std::mutex g_mu;
int g_enabled = -1;  // -1: from the environment at first use

std::map<std::tuple<int, int, int>, Entry>& cache()
{
    static auto* m = new std::map<std::tuple<int, int, int>, Entry>();
    return *m;
}

std::tuple<int, int, int> tup(const TuneKey& k)
{
    return std::make_tuple(k.kw, k.card_bucket, k.rows_bucket);
}

}  // namespace

TuneKey tune_key(int kw, uint64_t slots, int64_t rows)
{
    TuneKey k;
    k.kw = kw;
    int c = 0;
    while (c < 63 && ((uint64_t)1 << c) < slots) ++c;
    k.card_bucket = c;
    int r = 0;
    while (r < 62 && ((int64_t)2 << r) <= rows) ++r;
    k.rows_bucket = r;
    return k;
}

bool tune_lookup(const TuneKey& key, TuneConfig* cfg)
{
    std::lock_guard<std::mutex> lock(g_mu);
    auto it = cache().find(tup(key));
    if (it == cache().end()) return false;
    *cfg = it->second.cfg;
    return true;
}

void tune_store(const TuneKey& key, const TuneConfig& cfg, double ms)
{
    std::lock_guard<std::mutex> lock(g_mu);
    cache()[tup(key)] = Entry{cfg, ms};
}

bool tune_enabled()
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_enabled < 0) {
        const char* e = getenv("QDB_GPU_AUTOTUNE");
        g_enabled = (e && e[0] == '1') ? 1 : 0;
    }
    return g_enabled == 1;
}

bool tune_forced(TuneConfig* cfg)
{
    const char* e = getenv("QDB_GPU_TUNE_FORCE");
    if (!e || !*e) return false;
    TuneConfig c;
    c.part_blocks = 0;
    c.lds_copies = 1;
    const int got = sscanf(e, "%d,%d,%d,%d,%d,%d", &c.path, &c.lds_kib, &c.grid_per_cu, &c.fanout_shift, &c.part_blocks,
        &c.lds_copies);
    if (got < 4) return false;
    *cfg = c;
    return true;
}

std::vector<TuneConfig> tune_candidates(bool bucketed_ok)
{
    // LDS budget x workgroups per CU within the 160 KiB of a CU, and the
    // bucket count at 1/2x, 1x and 2x the sizing rule's.
    std::vector<TuneConfig> v = {
        {0, 32, 4, 0, 0, 1},
        {0, 64, 2, 0, 0, 1},
        {0, 128, 1, 0, 0, 1},
        {0, 64, 2, 0, 0, 8},   // eight copies of the LDS state words (a few groups)
    };
    if (bucketed_ok) {
        // and the partition at up to 2048 or 512 workgroups (fewer, longer
        // row ranges per workgroup: longer runs per bucket).
        for (int blocks : {2048, 512}) {
            v.push_back({1, 0, 0, -1, blocks, 1});
            v.push_back({1, 0, 0, 0, blocks, 1});
            v.push_back({1, 0, 0, 1, blocks, 1});
        }
    }
    return v;
}

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_autotune_enable(int on)
{
    const int was = tune_enabled() ? 1 : 0;
    std::lock_guard<std::mutex> lock(g_mu);
    g_enabled = on ? 1 : 0;
    return was;
}

extern "C" int64_t qdb_gpu_autotune_entries(QdbTuneEntry* out, int64_t max)
{
    std::lock_guard<std::mutex> lock(g_mu);
    int64_t n = 0;
    for (const auto& kv : cache()) {
        if (out && n < max) {
            QdbTuneEntry& e = out[n];
            memset(&e, 0, sizeof(e));
            e.KeyWords = std::get<0>(kv.first);
            e.CardBucket = std::get<1>(kv.first);
            e.RowsBucket = std::get<2>(kv.first);
            e.Path = kv.second.cfg.path;
            e.LdsKiB = kv.second.cfg.lds_kib;
            e.GridPerCu = kv.second.cfg.grid_per_cu;
            e.FanoutShift = kv.second.cfg.fanout_shift;
            e.PartBlocks = kv.second.cfg.part_blocks;
            e.LdsCopies = kv.second.cfg.lds_copies;
            e.Ms = kv.second.ms;
        }
        ++n;
    }
    return n;
}

extern "C" void qdb_gpu_autotune_clear(void)
{
    std::lock_guard<std::mutex> lock(g_mu);
    cache().clear();
}

extern "C" int qdb_gpu_autotune_save(const char* path)
{
    if (!path) return -1;
    std::lock_guard<std::mutex> lock(g_mu);
    FILE* f = fopen(path, "w");
    if (!f) return -1;
    fprintf(f, "# kw card_bucket rows_bucket path lds_kib grid_per_cu fanout_shift part_blocks lds_copies ms\n");
    for (const auto& kv : cache()) {
        const TuneConfig& c = kv.second.cfg;
        fprintf(f, "%d %d %d %d %d %d %d %d %d %.6f\n", std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), c.path,
            c.lds_kib, c.grid_per_cu, c.fanout_shift, c.part_blocks, c.lds_copies, kv.second.ms);
    }
    return fclose(f) == 0 ? 0 : -1;
}

extern "C" int64_t qdb_gpu_autotune_load(const char* path)
{
    if (!path) return -1;
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    char line[256];
    int64_t n = 0;
    while (fgets(line, sizeof(line), f)) {
        if (line[0] == '#') continue;
        TuneKey k;
        TuneConfig c;
        double ms = 0;
        if (sscanf(line, "%d %d %d %d %d %d %d %d %d %lf", &k.kw, &k.card_bucket, &k.rows_bucket, &c.path, &c.lds_kib,
                &c.grid_per_cu, &c.fanout_shift, &c.part_blocks, &c.lds_copies, &ms) != 10) {
            continue;
        }
        const bool valid = (c.path == 0 && (c.lds_kib == 32 || c.lds_kib == 64 || c.lds_kib == 128) && c.grid_per_cu >= 1 &&
                               c.grid_per_cu <= 8 && c.lds_copies >= 1 && c.lds_copies <= 16) ||
            (c.path == 1 && c.fanout_shift >= -2 && c.fanout_shift <= 2 && c.part_blocks >= 0 && c.part_blocks <= 2048);
        if (!valid) continue;
        tune_store(k, c, ms);
        ++n;
    }
    fclose(f);
    return n;
}
