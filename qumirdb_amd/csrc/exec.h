// Host operator processors over the C-ABI entry points, mirroring the
// reference's qdb/exec processors: TAggregateProcessor
// (qdb/exec/aggregate_exec.h:15-44, aggregate_exec.cpp:52-156) and
// TRowJoinProcessor (qdb/exec/row_join_exec.h:40-152,
// row_join_exec.cpp:22-276). They own the fixed-size state blobs, drive the
// op-code protocol (EnsureInit -> Update per batch -> Finish;
// Init -> Append* -> Build -> Probe* -> Materialize) and turn the kernels'
// return codes into std::runtime_error, as the reference does.
#pragma once

#include <array>
#include <cstdint>
#include <deque>
#include <functional>
#include <stdexcept>
#include <vector>

#include "../../include/qdb_gpu.h"

namespace NQdbGpu {

inline void Retain(TRowSet* rs)
{
    ++rs->RefCount;
}

inline void Release(TRowSet* rs)
{
    if (--rs->RefCount == 0 && rs->Destroy) rs->Destroy(rs);
}

enum class EAggregateTableLayout { Swiss, Row };

// TAggregateTablePlan (qdb/exec/plan_data.h:226-250).
struct TAggregateTablePlan {
    EAggregateTableLayout Layout = EAggregateTableLayout::Swiss;
    int64_t InitialCapacity = 8;
    int64_t InitialGroups = 0;
    int64_t RowSwitchGroups = 0;
    int64_t RowExpectedGroups = 0;
};

class TAggregateProcessor {
public:
    explicit TAggregateProcessor(QdbAggregateKernels kernels, TAggregateTablePlan plan = {});
    ~TAggregateProcessor();
    TAggregateProcessor(const TAggregateProcessor&) = delete;
    TAggregateProcessor& operator=(const TAggregateProcessor&) = delete;

    void Add(TRowSet& rowSet);
    // Updates many processors of one compiled aggregate, processors[i] with
    // rowSets[i], in one batched launch (qdb_gpu_agg_update_batched).
    static void AddBatched(TAggregateProcessor* const* processors, TRowSet* const* rowSets, int64_t n);
    bool Finish(TRowSet& rowSet);
    // Finishes many processors of one compiled aggregate, processors[i] into
    // rowSets[i], with one synchronization (qdb_gpu_agg_finish_batched).
    static void FinishBatched(TAggregateProcessor* const* processors, TRowSet* rowSets, int64_t n);
    int64_t Groups() const;
    bool Flush(TRowSet& rowSet);

private:
    void EnsureInit();

    QdbAggregateKernels Kernels_;
    TAggregateTablePlan Plan_;
    std::array<uint8_t, QDB_HASH_TABLE_SIZE> HashTable_ = {};
    bool Initialized_ = false;
    bool Finished_ = false;
};

enum class EJoinBuildSide { Auto, Left, Right };
enum class EJoinFetchResult { OK, NO_DATA, FINISHED };
enum class EJoinProcessorResult { OK, NEED_DATA, FINISHED };

class TRowJoinProcessor {
public:
    using TFetch = std::function<EJoinFetchResult(TRowSet&)>;

    // outputBatchRows: rows per materialized output batch (the reference uses
    // 16384, row_join_exec.cpp:14).
    TRowJoinProcessor(QdbRowJoinKernels kernels, EJoinBuildSide buildSide, size_t probeHoldBytes,
        int64_t outputBatchRows = 16384);
    ~TRowJoinProcessor();
    TRowJoinProcessor(const TRowJoinProcessor&) = delete;
    TRowJoinProcessor& operator=(const TRowJoinProcessor&) = delete;

    EJoinProcessorResult Process(const TFetch& left, const TFetch& right, TRowSet& output);
    EJoinBuildSide RequiredInputSide() const;

    // The explicit steps Process is made of.
    void AppendBuild(TRowSet& batch);
    void FinishBuild();
    // Probes the batch and queues its materialized output; releases it.
    void ProbeBatch(TRowSet& batch);
    bool DrainReadyOutput(TRowSet& rowSet);
    // Probes without materializing: the pairs stay in the pair buffer until
    // the next probe (for fused aggregates and pair checks).
    int64_t ProbeOnly(TRowSet& batch);
    const QdbPairBuffer& Pairs() const { return PairBuffer_; }
    void* ArenaBlob() { return Arena_.data(); }

private:
    void EnsureInit();
    bool Dispatch(int64_t op, TRowSet* batch);
    bool PullBuildBatch(const TFetch& build);
    bool PullProbeBatch(const TFetch& probe);
    bool CanHoldProbe() const;
    uint64_t BatchBytes(const TRowSet& batch) const;

    QdbRowJoinKernels Kernels_;
    EJoinBuildSide BuildSide_;
    size_t ProbeHoldBytes_;
    int64_t OutputBatchRows_;
    std::array<uint8_t, QDB_ROW_ARENA_SIZE> Arena_ = {};
    std::array<uint8_t, QDB_ROW_TABLE_SIZE> Table_ = {};
    std::array<uint8_t, QDB_ROW_SCRATCH_SIZE> Scratch_ = {};
    QdbPairBuffer PairBuffer_ = {0, 0, nullptr};
    struct THeldProbe {
        TRowSet Batch;
        uint64_t Bytes;
    };
    std::deque<THeldProbe> PendingProbe_;
    uint64_t HeldProbeBytes_ = 0;
    std::deque<TRowSet> ReadyOutput_;
    bool Initialized_ = false;
    bool BuildDone_ = false;
    bool ProbeDone_ = false;
};

}  // namespace NQdbGpu
