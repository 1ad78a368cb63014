// Host drivers of the device operators: they own the fixed-size state blobs
// of one aggregate or one row join, drive the kernels' op-code protocol
// (qdb/kernel/compiler.h:217-230, 363-383) and check the invariants the
// reference's callers check, so that a caller holding only the C-ABI gets the
// same behaviour the reference's processors give its kernels:
//
//   aggregate  Init (or InitRows) once, before the first batch, with the
//              plan's capacity and its row-layout hints
//              (qdb/exec/aggregate_exec.cpp:67-95); Update per batch (< 0 is
//              an error, :107-113); at the end the table's Size is read from
//              the blob BEFORE agg_finish_rowset, the finalized row count
//              must equal it (:128-132), and the kernel-owned output is
//              adopted (qdb/exec/kernel_rowset.h:30-38).
//   row join   Init, Append per build batch, Build once, then Probe per probe
//              batch; the probe's pairs are materialized in output batches
//              of a fixed row count before the probe batch is released, and
//              the pair count is reset after each drain
//              (qdb/exec/row_join_exec.cpp:167-190).
//
// The drivers are push-style: a batch is handed in by the caller and stays
// the caller's (borrowed, never released here); outputs queue until the
// caller takes them. The reference's pull loop (fetch callbacks, held probe
// batches, the build-side choice) is scheduler work and has no counterpart.
#pragma once

#include <array>
#include <cstdint>
#include <deque>
#include <memory>
#include <stdexcept>
#include <vector>

#include "../../include/qdb_gpu.h"

namespace NQdbGpu {

// A driver error carries the library's error text.
class TDriverError : public std::runtime_error {
public:
    explicit TDriverError(const char* what);
};

// Takes ownership of a kernel-owned output: every owner is detached from the
// host heap (qdb_gpu_set_host_heap's detach, TKernelHeap::Detach in the
// reference) and the destroy callback is installed.
void AdoptOutput(TRowSet& rowSet);

// The table plan the reference's planner hands to its processor
// (TAggregateTablePlan, qdb/exec/plan_data.h:226-250).
struct TAggregatePlan {
    bool RowLayout = false;
    int64_t InitialCapacity = 8;   // Init's argument: a power of two, at least 8
    int64_t InitialGroups = 0;     // InitRows' argument
    int64_t RowSwitchGroups = 0;
    int64_t RowExpectedGroups = 0;
};

class TAggregateDriver {
public:
    TAggregateDriver(const QdbAggregateKernels& kernels, const TAggregatePlan& plan);
    ~TAggregateDriver();
    TAggregateDriver(const TAggregateDriver&) = delete;
    TAggregateDriver& operator=(const TAggregateDriver&) = delete;

    void Consume(const TRowSet& batch);
    // drivers[i] consumes batches[i], every driver of one compiled
    // aggregate, in one batched launch.
    static void ConsumeMany(TAggregateDriver* const* drivers, const TRowSet* const* batches, int64_t n);
    // Finalizes into `out` (adopted); returns the row count.
    int64_t Emit(TRowSet& out);
    // drivers[i] finalizes into outs[i], with one synchronization.
    static void EmitMany(TAggregateDriver* const* drivers, TRowSet* outs, int64_t n);
    // Groups so far (the table's Size); 0 before the first batch and after Emit.
    int64_t Groups() const;

private:
    enum class EState { Fresh, Open, Emitted };
    void Open();
    int64_t TableSize() const;
    void* Blob() { return Table_.data(); }

    QdbAggregateKernels Kernels_;
    TAggregatePlan Plan_;
    std::array<uint8_t, QDB_HASH_TABLE_SIZE> Table_ = {};
    EState State_ = EState::Fresh;
};

// The bounded, adaptive partial aggregate (TPartialAggregatePolicy,
// qdb/exec/plan_data.h:234-250; behaviour of TPartialAggregateProcessor,
// qdb/exec/aggregate_exec.cpp:158-233): a table that emits its groups and
// restarts when it reaches MaxGroups, and that, once SampleRows input rows
// have been hashed without an average of MinReduction rows per group
// (emitted groups included), emits its table and passes every later batch
// through as single-row partial states (qdb_gpu_agg_partial_states). The
// decision is one-way.
struct TPartialPolicy {
    uint64_t SampleRows = 128 * 1024;
    uint64_t MinReduction = 8;
    int64_t MaxGroups = 1 << 20;
};

class TPartialAggregateDriver {
public:
    TPartialAggregateDriver(const QdbAggregateKernels& kernels, const TAggregatePlan& plan, const TPartialPolicy& policy,
        bool passThrough);
    // Consumes a batch; an output may then be ready (Next).
    void Consume(const TRowSet& batch);
    // The pending output (adopted); false when none.
    bool Next(TRowSet& out);
    // The table's remaining groups; false when rows are passed through (the
    // table was emitted when that began).
    bool Finish(TRowSet& out);
    bool PassingThrough() const { return PassingThrough_; }

private:
    void Flush();

    QdbAggregateKernels Kernels_;
    TAggregatePlan Plan_;
    TPartialPolicy Policy_;
    bool PassThrough_;
    std::unique_ptr<TAggregateDriver> Table_;
    std::deque<TRowSet> Pending_;
    uint64_t HashedRows_ = 0;
    uint64_t EmittedGroups_ = 0;
    bool PassingThrough_ = false;
};

class TJoinDriver {
public:
    TJoinDriver(const QdbRowJoinKernels& kernels, int64_t outputRows);
    ~TJoinDriver();
    TJoinDriver(const TJoinDriver&) = delete;
    TJoinDriver& operator=(const TJoinDriver&) = delete;

    void AddBuild(const TRowSet& batch);
    void Seal();
    // Probes the batch; with `materialize` its matches are gathered into
    // queued output batches and the pairs are drained, else the pairs stay
    // readable (Pairs()) until the next probe. Returns the pair count.
    int64_t Probe(const TRowSet& batch, bool materialize);
    // The next queued output batch (adopted); false when none is left.
    bool Next(TRowSet& out);
    const QdbPairBuffer& Pairs() const { return Pairs_; }
    void* Arena() { return Arena_.data(); }

    // Seal() of every driver (one compiled join) with one synchronisation
    // for all their dense tables (qdb_gpu_join_build_batched).
    static void SealMany(TJoinDriver* const* drivers, int64_t n);
    // Probe(batches[i], false) of every driver in one persistent launch
    // (qdb_gpu_join_probe_batched); each driver's pairs stay readable.
    static void ProbeMany(TJoinDriver* const* drivers, const TRowSet* const* batches, int64_t n);

private:
    bool Op(int64_t op, const TRowSet* batch);

    QdbRowJoinKernels Kernels_;
    int64_t OutputRows_;
    std::array<uint8_t, QDB_ROW_ARENA_SIZE> Arena_ = {};
    std::array<uint8_t, QDB_ROW_TABLE_SIZE> Table_ = {};
    std::array<uint8_t, QDB_ROW_SCRATCH_SIZE> Scratch_ = {};
    QdbPairBuffer Pairs_ = {0, 0, nullptr};
    std::deque<TRowSet> Outputs_;
    bool Sealed_ = false;
};

}  // namespace NQdbGpu
