// Host operator processors (see exec.h) and their C exports for callers that
// bind by C-ABI (tests, other host languages).
#include "exec.h"

#include <string.h>

#include <memory>
#include <string>

namespace NQdbGpu {

namespace {

// Mirrors the first fields of the HashTable blob: the processor reads Size
// (qdb/exec/aggregate_exec.cpp:18-39).
struct THashTableHead {
    void* Words[8];
    int64_t Capacity;
    int64_t Size;
};

constexpr int64_t InitialCapacity = 8;

std::runtime_error Error(const char* what)
{
    std::string text = what;
    const char* detail = qdb_gpu_last_error();
    if (detail && *detail) {
        text += ": ";
        text += detail;
    }
    return std::runtime_error(text);
}

uint64_t BitCeil(uint64_t v)
{
    uint64_t c = 1;
    while (c < v) c <<= 1;
    return c;
}

}  // namespace

TAggregateProcessor::TAggregateProcessor(QdbAggregateKernels kernels, TAggregateTablePlan plan)
    : Kernels_(kernels)
    , Plan_(plan)
{
    // aht_init only accepts a power-of-two capacity of at least 8.
    Plan_.InitialCapacity = (int64_t)BitCeil((uint64_t)(Plan_.InitialCapacity > InitialCapacity ? Plan_.InitialCapacity : InitialCapacity));
    if (Kernels_.RowBytes == 0
        && (Plan_.Layout == EAggregateTableLayout::Row || Plan_.RowSwitchGroups != 0 || Plan_.RowExpectedGroups != 0)) {
        throw std::invalid_argument("aggregate has no row layout");
    }
}

TAggregateProcessor::~TAggregateProcessor()
{
    if (Initialized_ && !Finished_) qdb_gpu_aggregate_destroy(HashTable_.data());
}

void TAggregateProcessor::EnsureInit()
{
    if (Initialized_) return;
    void* table = HashTable_.data();
    const bool initialized = Plan_.Layout == EAggregateTableLayout::Row
        ? Kernels_.Dispatch(table, nullptr, Plan_.InitialGroups, QDB_AGG_OP_INIT_ROWS) != 0
        : Kernels_.Dispatch(table, nullptr, Plan_.InitialCapacity, QDB_AGG_OP_INIT) != 0;
    if (!initialized) throw Error("aggregate initialization failed");
    Initialized_ = true;
    if (Plan_.RowSwitchGroups != 0) Kernels_.Dispatch(table, nullptr, Plan_.RowSwitchGroups, QDB_AGG_OP_SWITCH_TO_ROWS_AT);
    if (Plan_.RowExpectedGroups != 0) Kernels_.Dispatch(table, nullptr, Plan_.RowExpectedGroups, QDB_AGG_OP_EXPECT_ROW_GROUPS);
}

void TAggregateProcessor::Add(TRowSet& rowSet)
{
    if (Finished_) throw std::runtime_error("aggregate processor is already finished");
    EnsureInit();
    if (Kernels_.Dispatch(HashTable_.data(), &rowSet, 0, QDB_AGG_OP_UPDATE) < 0) throw Error("aggregate update failed");
}

void TAggregateProcessor::AddBatched(TAggregateProcessor* const* processors, TRowSet* const* rowSets, int64_t n)
{
    if (n <= 0) return;
    std::vector<void*> tables((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        TAggregateProcessor* p = processors[i];
        if (p->Finished_) throw std::runtime_error("aggregate processor is already finished");
        if (p->Kernels_.Slot != processors[0]->Kernels_.Slot) throw std::invalid_argument("batched processors of different aggregates");
        p->EnsureInit();
        tables[(size_t)i] = p->HashTable_.data();
    }
    if (qdb_gpu_agg_update_batched(&processors[0]->Kernels_, tables.data(), rowSets, n) != 0) throw Error("batched aggregate update failed");
}

void TAggregateProcessor::FinishBatched(TAggregateProcessor* const* processors, TRowSet* rowSets, int64_t n)
{
    if (n <= 0) return;
    std::vector<void*> tables((size_t)n);
    std::vector<int64_t> sizes((size_t)n), rows((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        TAggregateProcessor* p = processors[i];
        if (p->Finished_) throw std::runtime_error("aggregate processor is already finished");
        if (p->Kernels_.Slot != processors[0]->Kernels_.Slot) throw std::invalid_argument("batched processors of different aggregates");
        p->EnsureInit();
        tables[(size_t)i] = p->HashTable_.data();
        sizes[(size_t)i] = reinterpret_cast<THashTableHead*>(p->HashTable_.data())->Size;
    }
    for (int64_t i = 0; i < n; ++i) processors[i]->Finished_ = true;
    if (qdb_gpu_agg_finish_batched(&processors[0]->Kernels_, tables.data(), rowSets, rows.data(), n) != 0) {
        throw Error("batched aggregate finalize failed");
    }
    for (int64_t i = 0; i < n; ++i) {
        // A global aggregate over an empty input emits its one group.
        if (rows[(size_t)i] != sizes[(size_t)i] && !(sizes[(size_t)i] == 0 && rows[(size_t)i] == 1)) {
            for (int64_t k = 0; k < n; ++k) {
                if (rowSets[k].Destroy) rowSets[k].Destroy(&rowSets[k]);
            }
            throw std::runtime_error("aggregate finalize returned an unexpected row count");
        }
    }
}

bool TAggregateProcessor::Finish(TRowSet& rowSet)
{
    if (Finished_) return false;
    Finished_ = true;
    EnsureInit();
    const int64_t size = reinterpret_cast<THashTableHead*>(HashTable_.data())->Size;
    const int64_t finalized = Kernels_.FinishRowSet(HashTable_.data(), &rowSet);
    // A global aggregate over an empty input emits its one group.
    if (finalized < 0) throw Error("aggregate finalize failed");
    if (finalized != size && !(size == 0 && finalized == 1)) {
        throw std::runtime_error("aggregate finalize returned an unexpected row count");
    }
    return true;
}

int64_t TAggregateProcessor::Groups() const
{
    if (!Initialized_ || Finished_) return 0;
    return reinterpret_cast<const THashTableHead*>(HashTable_.data())->Size;
}

bool TAggregateProcessor::Flush(TRowSet& rowSet)
{
    if (Groups() == 0) return false;
    Finish(rowSet);
    Initialized_ = false;
    Finished_ = false;
    return true;
}

TRowJoinProcessor::TRowJoinProcessor(QdbRowJoinKernels kernels, EJoinBuildSide buildSide, size_t probeHoldBytes,
    int64_t outputBatchRows)
    : Kernels_(kernels)
    , BuildSide_(buildSide)
    , ProbeHoldBytes_(probeHoldBytes)
    , OutputBatchRows_(outputBatchRows)
{
    if (BuildSide_ == EJoinBuildSide::Auto) throw std::runtime_error("row join requires a fixed build side");
}

TRowJoinProcessor::~TRowJoinProcessor()
{
    while (!ReadyOutput_.empty()) {
        Release(&ReadyOutput_.front());
        ReadyOutput_.pop_front();
    }
    for (auto& held : PendingProbe_) Release(&held.Batch);
    qdb_gpu_join_destroy(Arena_.data(), Table_.data(), Scratch_.data(), &PairBuffer_);
}

bool TRowJoinProcessor::Dispatch(int64_t op, TRowSet* batch)
{
    return Kernels_.Dispatch(Arena_.data(), Table_.data(), Scratch_.data(), &PairBuffer_, batch, nullptr, op);
}

void TRowJoinProcessor::EnsureInit()
{
    if (Initialized_) return;
    Initialized_ = true;
    TRowSet none{};
    if (!Dispatch(QDB_JR_OP_INIT, &none)) throw Error("row join initialization failed");
}

bool TRowJoinProcessor::CanHoldProbe() const
{
    return HeldProbeBytes_ < ProbeHoldBytes_;
}

uint64_t TRowJoinProcessor::BatchBytes(const TRowSet& batch) const
{
    // All rows, no selection: the width of every fixed column.
    return (uint64_t)batch.RowCount * (uint64_t)(batch.ColumnCount > 0 ? batch.ColumnCount : 1) * 8u;
}

void TRowJoinProcessor::AppendBuild(TRowSet& batch)
{
    EnsureInit();
    if (BuildDone_) throw std::runtime_error("row join append after build");
    if (!Dispatch(QDB_JR_OP_APPEND, &batch)) {
        Release(&batch);
        throw Error("row join append failed");
    }
    Release(&batch);
}

void TRowJoinProcessor::FinishBuild()
{
    EnsureInit();
    TRowSet none{};
    if (!Dispatch(QDB_JR_OP_BUILD, &none)) throw Error("row join table build failed");
    BuildDone_ = true;
}

int64_t TRowJoinProcessor::ProbeOnly(TRowSet& batch)
{
    if (!BuildDone_) throw std::runtime_error("row join probe before build");
    if (!Dispatch(QDB_JR_OP_PROBE, &batch)) throw Error("row join probe failed");
    return PairBuffer_.Count;
}

void TRowJoinProcessor::ProbeBatch(TRowSet& batch)
{
    if (!Dispatch(QDB_JR_OP_PROBE, &batch)) {
        Release(&batch);
        throw Error("row join probe failed");
    }
    // The pairs refer to this batch, so its matches are materialized before
    // it is released.
    int64_t cursor = 0;
    while (cursor < PairBuffer_.Count) {
        TRowSet out{};
        const int64_t produced = Kernels_.Materialize(&PairBuffer_, &batch, cursor, OutputBatchRows_, &out);
        if (produced <= 0) {
            Release(&batch);
            throw Error("row join materialize failed");
        }
        ReadyOutput_.push_back(out);
        cursor += produced;
    }
    PairBuffer_.Count = 0;
    Release(&batch);
}

bool TRowJoinProcessor::DrainReadyOutput(TRowSet& rowSet)
{
    if (ReadyOutput_.empty()) return false;
    rowSet = ReadyOutput_.front();
    ReadyOutput_.pop_front();
    return true;
}

bool TRowJoinProcessor::PullBuildBatch(const TFetch& build)
{
    TRowSet batch{};
    const auto fetch = build(batch);
    if (fetch == EJoinFetchResult::NO_DATA) return false;
    if (fetch == EJoinFetchResult::FINISHED) {
        FinishBuild();
        return true;
    }
    AppendBuild(batch);
    return true;
}

bool TRowJoinProcessor::PullProbeBatch(const TFetch& probe)
{
    TRowSet batch{};
    const auto fetch = probe(batch);
    if (fetch == EJoinFetchResult::NO_DATA) return false;
    if (fetch == EJoinFetchResult::FINISHED) {
        ProbeDone_ = true;
        return true;
    }
    if (!BuildDone_) {
        const auto bytes = BatchBytes(batch);
        HeldProbeBytes_ += bytes;
        PendingProbe_.push_back({batch, bytes});
        return true;
    }
    ProbeBatch(batch);
    return true;
}

EJoinProcessorResult TRowJoinProcessor::Process(const TFetch& left, const TFetch& right, TRowSet& output)
{
    EnsureInit();
    if (DrainReadyOutput(output)) return EJoinProcessorResult::OK;
    const TFetch& build = BuildSide_ == EJoinBuildSide::Left ? left : right;
    const TFetch& probe = BuildSide_ == EJoinBuildSide::Left ? right : left;
    if (!BuildDone_) {
        if (!PullBuildBatch(build) && !ProbeDone_ && CanHoldProbe()) PullProbeBatch(probe);
        if (!BuildDone_) return EJoinProcessorResult::NEED_DATA;
    }
    while (!PendingProbe_.empty()) {
        TRowSet batch = PendingProbe_.front().Batch;
        HeldProbeBytes_ -= PendingProbe_.front().Bytes;
        PendingProbe_.pop_front();
        ProbeBatch(batch);
        if (DrainReadyOutput(output)) return EJoinProcessorResult::OK;
    }
    if (ProbeDone_) return EJoinProcessorResult::FINISHED;
    if (!PullProbeBatch(probe)) return EJoinProcessorResult::NEED_DATA;
    if (DrainReadyOutput(output)) return EJoinProcessorResult::OK;
    return ProbeDone_ ? EJoinProcessorResult::FINISHED : EJoinProcessorResult::NEED_DATA;
}

EJoinBuildSide TRowJoinProcessor::RequiredInputSide() const
{
    if (!BuildDone_) return ProbeDone_ || !CanHoldProbe() ? BuildSide_ : EJoinBuildSide::Auto;
    return BuildSide_ == EJoinBuildSide::Left ? EJoinBuildSide::Right : EJoinBuildSide::Left;
}

}  // namespace NQdbGpu

// ---------------------------------------------------------------------------
// C exports of the processors. Errors are reported through the return value
// and qdb_gpu_last_error-style text in qdb_gpu_exec_last_error.
// ---------------------------------------------------------------------------

namespace {

thread_local std::string g_exec_error;

template <typename F>
int Guard(F&& f)
{
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_exec_error = e.what();
        return -1;
    }
}

// A batch handed to a processor is consumed (Released). The C callers keep
// ownership, so batches are passed as retained views.
TRowSet View(const TRowSet* rs)
{
    TRowSet v = *rs;
    v.Destroy = nullptr;
    v.RefCount = 1;
    return v;
}

}  // namespace

extern "C" {

const char* qdb_gpu_exec_last_error(void)
{
    return g_exec_error.c_str();
}

void* qdb_gpu_agg_processor_new(const QdbAggregateKernels* kernels, int64_t initial_capacity, int row_layout,
    int64_t initial_groups)
{
    NQdbGpu::TAggregateProcessor* p = nullptr;
    Guard([&] {
        NQdbGpu::TAggregateTablePlan plan;
        plan.InitialCapacity = initial_capacity;
        plan.Layout = row_layout ? NQdbGpu::EAggregateTableLayout::Row : NQdbGpu::EAggregateTableLayout::Swiss;
        plan.InitialGroups = initial_groups;
        p = new NQdbGpu::TAggregateProcessor(*kernels, plan);
    });
    return p;
}

int qdb_gpu_agg_processor_add(void* p, const TRowSet* batch)
{
    return Guard([&] {
        TRowSet v = View(batch);
        static_cast<NQdbGpu::TAggregateProcessor*>(p)->Add(v);
    });
}

int qdb_gpu_agg_processor_add_batched(void* const* procs, const TRowSet* const* batches, int64_t n)
{
    return Guard([&] {
        std::vector<TRowSet> views((size_t)(n > 0 ? n : 0));
        std::vector<TRowSet*> ptrs((size_t)(n > 0 ? n : 0));
        for (int64_t i = 0; i < n; ++i) {
            views[(size_t)i] = View(batches[i]);
            ptrs[(size_t)i] = &views[(size_t)i];
        }
        NQdbGpu::TAggregateProcessor::AddBatched(reinterpret_cast<NQdbGpu::TAggregateProcessor* const*>(procs), ptrs.data(), n);
    });
}

int64_t qdb_gpu_agg_processor_finish(void* p, TRowSet* out)
{
    int64_t rows = -1;
    if (Guard([&] {
            memset(out, 0, sizeof(*out));
            static_cast<NQdbGpu::TAggregateProcessor*>(p)->Finish(*out);
            rows = out->RowCount;
        }) != 0) {
        return -1;
    }
    return rows;
}

int qdb_gpu_agg_processor_finish_batched(void* const* procs, TRowSet* outs, int64_t n)
{
    return Guard([&] {
        NQdbGpu::TAggregateProcessor::FinishBatched(reinterpret_cast<NQdbGpu::TAggregateProcessor* const*>(procs), outs, n);
    });
}

int64_t qdb_gpu_agg_processor_groups(void* p)
{
    return static_cast<NQdbGpu::TAggregateProcessor*>(p)->Groups();
}

void qdb_gpu_agg_processor_free(void* p)
{
    delete static_cast<NQdbGpu::TAggregateProcessor*>(p);
}

void* qdb_gpu_join_processor_new(const QdbRowJoinKernels* kernels, int build_is_left, int64_t output_batch_rows)
{
    NQdbGpu::TRowJoinProcessor* p = nullptr;
    Guard([&] {
        p = new NQdbGpu::TRowJoinProcessor(*kernels, build_is_left ? NQdbGpu::EJoinBuildSide::Left : NQdbGpu::EJoinBuildSide::Right,
            16u << 20, output_batch_rows);
    });
    return p;
}

int qdb_gpu_join_processor_append(void* p, const TRowSet* batch)
{
    return Guard([&] {
        TRowSet v = View(batch);
        static_cast<NQdbGpu::TRowJoinProcessor*>(p)->AppendBuild(v);
    });
}

int qdb_gpu_join_processor_build(void* p)
{
    return Guard([&] { static_cast<NQdbGpu::TRowJoinProcessor*>(p)->FinishBuild(); });
}

// Probes and materializes; returns the number of output batches queued.
int64_t qdb_gpu_join_processor_probe(void* p, const TRowSet* batch)
{
    int64_t before = 0;
    NQdbGpu::TRowJoinProcessor* j = static_cast<NQdbGpu::TRowJoinProcessor*>(p);
    if (Guard([&] {
            TRowSet v = View(batch);
            j->ProbeBatch(v);
        }) != 0) {
        return -1;
    }
    return before;
}

// Probes only; returns the pair count (pairs readable with
// qdb_gpu_join_processor_pairs until the next probe).
int64_t qdb_gpu_join_processor_probe_only(void* p, const TRowSet* batch)
{
    int64_t n = -1;
    NQdbGpu::TRowJoinProcessor* j = static_cast<NQdbGpu::TRowJoinProcessor*>(p);
    if (Guard([&] {
            TRowSet v = View(batch);
            n = j->ProbeOnly(v);
        }) != 0) {
        return -1;
    }
    return n;
}

const QdbPairBuffer* qdb_gpu_join_processor_pairs(void* p)
{
    return &static_cast<NQdbGpu::TRowJoinProcessor*>(p)->Pairs();
}

void* qdb_gpu_join_processor_arena(void* p)
{
    return static_cast<NQdbGpu::TRowJoinProcessor*>(p)->ArenaBlob();
}

// Next queued output batch; 1 when one was returned, 0 when none is left.
int qdb_gpu_join_processor_next(void* p, TRowSet* out)
{
    return static_cast<NQdbGpu::TRowJoinProcessor*>(p)->DrainReadyOutput(*out) ? 1 : 0;
}

void qdb_gpu_join_processor_free(void* p)
{
    delete static_cast<NQdbGpu::TRowJoinProcessor*>(p);
}

void qdb_gpu_rowset_release(TRowSet* rs)
{
    if (rs && rs->Destroy) rs->Destroy(rs);
}

}  // extern "C"
