// The aggregate autotuner's cache: one launch configuration per shape
// bucket (key width, cardinality bucket, row-count bucket), chosen by timing
// the candidates on the first batch of that shape and reused afterwards
// (optionally persisted to a text file between processes).
#pragma once

#include <stdint.h>

#include <vector>

namespace qdbgpu {

// How one aggregate update runs.
struct TuneConfig {
    int32_t path;          // 0: LDS tables + HBM table (one pass), 1: bucketed (partition, LDS aggregate, merge)
    int32_t lds_kib;       // path 0: LDS table budget of a workgroup, KiB
    int32_t grid_per_cu;   // path 0: workgroups per CU
    int32_t fanout_shift;  // path 1: log2 of the bucket count relative to the sizing rule
    int32_t part_blocks;   // path 1: most partition workgroups (0: the sizing rule's 2048)
    int32_t lds_copies;    // path 0: copies of the LDS state words (1: none)
};

// A shape bucket: key words, ceil(log2(table slots)), floor(log2(rows)).
struct TuneKey {
    int32_t kw;
    int32_t card_bucket;
    int32_t rows_bucket;
};

TuneKey tune_key(int kw, uint64_t slots, int64_t rows);
// The configuration cached for `key`; false when none is.
bool tune_lookup(const TuneKey& key, TuneConfig* cfg);
void tune_store(const TuneKey& key, const TuneConfig& cfg, double ms);
// Whether updates of an uncached shape are tuned first (qdb_gpu_autotune_enable,
// or QDB_GPU_AUTOTUNE=1 in the environment).
bool tune_enabled();
// A configuration forced for every update (QDB_GPU_TUNE_FORCE =
// "path,lds_kib,grid_per_cu,fanout_shift[,part_blocks[,lds_copies]]"); false
// when none is.
bool tune_forced(TuneConfig* cfg);
// The candidates for a shape: path-0 LDS budgets and grids, and path-1
// fan-outs when the bucketed path can serve.
std::vector<TuneConfig> tune_candidates(bool bucketed_ok);

}  // namespace qdbgpu
