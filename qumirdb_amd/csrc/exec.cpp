// Host drivers of the device operators (exec.h) and their C exports.
#include "exec.h"

#include <stddef.h>
#include <string.h>

#include <string>

namespace NQdbGpu {

namespace {

// The first fields of the HashTable blob (qdb/exec/aggregate_exec.cpp:18-39):
// the driver reads Size, at byte offset 72.
struct THashTableHead {
    void* Pointers[6];
    int64_t OwnedBlockCount;
    int64_t OwnedBlockCapacity;
    int64_t Capacity;
    int64_t Size;
};
static_assert(offsetof(THashTableHead, Size) == 72, "HashTable blob: Size at offset 72");

std::string WithDetail(const char* what)
{
    std::string text = what;
    const char* detail = qdb_gpu_last_error();
    if (detail && *detail) {
        text += ": ";
        text += detail;
    }
    return text;
}

int64_t PowerOfTwoAtLeast(int64_t v, int64_t floor)
{
    int64_t c = floor;
    while (c < v) c <<= 1;
    return c;
}

void ReleaseOutput(TRowSet& rs)
{
    if (rs.Destroy) rs.Destroy(&rs);
    else qdb_gpu_destroy_kernel_rowset(&rs);
    memset(&rs, 0, sizeof(rs));
}

}  // namespace

TDriverError::TDriverError(const char* what)
    : std::runtime_error(WithDetail(what))
{
}

void AdoptOutput(TRowSet& rowSet)
{
    qdb_gpu_adopt_kernel_rowset(&rowSet);
}

// ---------------------------------------------------------------------------
// Aggregate
// ---------------------------------------------------------------------------

TAggregateDriver::TAggregateDriver(const QdbAggregateKernels& kernels, const TAggregatePlan& plan)
    : Kernels_(kernels)
    , Plan_(plan)
{
    // The swiss Init takes a power of two of at least 8 slots.
    Plan_.InitialCapacity = PowerOfTwoAtLeast(Plan_.InitialCapacity, 8);
    const bool hints = Plan_.RowLayout || Plan_.RowSwitchGroups != 0 || Plan_.RowExpectedGroups != 0;
    if (hints && Kernels_.RowBytes == 0) throw std::invalid_argument("aggregate has no row layout");
}

TAggregateDriver::~TAggregateDriver()
{
    if (State_ == EState::Open) qdb_gpu_aggregate_destroy(Blob());
}

void TAggregateDriver::Open()
{
    if (State_ != EState::Fresh) return;
    const int64_t ok = Plan_.RowLayout
        ? Kernels_.Dispatch(Blob(), nullptr, Plan_.InitialGroups, QDB_AGG_OP_INIT_ROWS)
        : Kernels_.Dispatch(Blob(), nullptr, Plan_.InitialCapacity, QDB_AGG_OP_INIT);
    if (ok == 0) throw TDriverError("aggregate initialization failed");
    State_ = EState::Open;
    // Hints only set fields of the table.
    if (Plan_.RowSwitchGroups != 0) Kernels_.Dispatch(Blob(), nullptr, Plan_.RowSwitchGroups, QDB_AGG_OP_SWITCH_TO_ROWS_AT);
    if (Plan_.RowExpectedGroups != 0) Kernels_.Dispatch(Blob(), nullptr, Plan_.RowExpectedGroups, QDB_AGG_OP_EXPECT_ROW_GROUPS);
}

int64_t TAggregateDriver::TableSize() const
{
    return reinterpret_cast<const THashTableHead*>(Table_.data())->Size;
}

void TAggregateDriver::Consume(const TRowSet& batch)
{
    if (State_ == EState::Emitted) throw std::logic_error("aggregate already emitted");
    Open();
    TRowSet view = batch;
    if (Kernels_.Dispatch(Blob(), &view, 0, QDB_AGG_OP_UPDATE) < 0) throw TDriverError("aggregate update failed");
}

void TAggregateDriver::ConsumeMany(TAggregateDriver* const* drivers, const TRowSet* const* batches, int64_t n)
{
    if (n <= 0) return;
    std::vector<void*> tables((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        TAggregateDriver* d = drivers[i];
        if (d->State_ == EState::Emitted) throw std::logic_error("aggregate already emitted");
        if (d->Kernels_.Slot != drivers[0]->Kernels_.Slot) throw std::invalid_argument("batched drivers of different aggregates");
        d->Open();
        tables[(size_t)i] = d->Blob();
    }
    if (qdb_gpu_agg_update_batched(&drivers[0]->Kernels_, tables.data(), batches, n) != 0) {
        throw TDriverError("batched aggregate update failed");
    }
}

int64_t TAggregateDriver::Emit(TRowSet& out)
{
    if (State_ == EState::Emitted) throw std::logic_error("aggregate already emitted");
    Open();
    // The reference reads Size before finalizing and requires the row count
    // to equal it: empty-input groups count from Init on.
    const int64_t size = TableSize();
    State_ = EState::Emitted;
    memset(&out, 0, sizeof(out));
    const int64_t rows = Kernels_.FinishRowSet(Blob(), &out);
    if (rows < 0) {
        ReleaseOutput(out);
        throw TDriverError("aggregate finalize failed");
    }
    if (rows != size) {
        ReleaseOutput(out);
        throw std::runtime_error("aggregate finalize returned an unexpected row count");
    }
    AdoptOutput(out);
    return rows;
}

void TAggregateDriver::EmitMany(TAggregateDriver* const* drivers, TRowSet* outs, int64_t n)
{
    if (n <= 0) return;
    std::vector<void*> tables((size_t)n);
    std::vector<int64_t> sizes((size_t)n), rows((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        TAggregateDriver* d = drivers[i];
        if (d->State_ == EState::Emitted) throw std::logic_error("aggregate already emitted");
        if (d->Kernels_.Slot != drivers[0]->Kernels_.Slot) throw std::invalid_argument("batched drivers of different aggregates");
        d->Open();
        tables[(size_t)i] = d->Blob();
        sizes[(size_t)i] = d->TableSize();
    }
    for (int64_t i = 0; i < n; ++i) drivers[i]->State_ = EState::Emitted;
    if (qdb_gpu_agg_finish_batched(&drivers[0]->Kernels_, tables.data(), outs, rows.data(), n) != 0) {
        throw TDriverError("batched aggregate finalize failed");
    }
    bool counts_ok = true;
    for (int64_t i = 0; i < n; ++i) counts_ok = counts_ok && rows[(size_t)i] == sizes[(size_t)i];
    for (int64_t i = 0; i < n; ++i) {
        if (counts_ok) AdoptOutput(outs[i]);
        else ReleaseOutput(outs[i]);
    }
    if (!counts_ok) throw std::runtime_error("aggregate finalize returned an unexpected row count");
}

int64_t TAggregateDriver::Groups() const
{
    return State_ == EState::Open ? TableSize() : 0;
}

// ---------------------------------------------------------------------------
// Partial aggregate
// ---------------------------------------------------------------------------

TPartialAggregateDriver::TPartialAggregateDriver(const QdbAggregateKernels& kernels, const TAggregatePlan& plan,
    const TPartialPolicy& policy, bool passThrough)
    : Kernels_(kernels)
    , Plan_(plan)
    , Policy_(policy)
    , PassThrough_(passThrough)
    , Table_(new TAggregateDriver(kernels, plan))
{
    // An aggregate whose partial-state projection cannot be built keeps the
    // bounded table and never passes rows through (the results are the same).
    if (PassThrough_) {
        TRowSet empty{};
        TRowSet probe{};
        PassThrough_ = qdb_gpu_agg_partial_states(&Kernels_, &empty, &probe) == 0;
        if (PassThrough_) qdb_gpu_destroy_kernel_rowset(&probe);
    }
}

void TPartialAggregateDriver::Flush()
{
    // Emits the table's groups (when it has any) and starts a new table.
    if (Table_->Groups() > 0) {
        TRowSet out{};
        Table_->Emit(out);
        Pending_.push_back(out);
    }
    Table_.reset(new TAggregateDriver(Kernels_, Plan_));
}

void TPartialAggregateDriver::Consume(const TRowSet& batch)
{
    if (PassingThrough_) {
        TRowSet states{};
        if (qdb_gpu_agg_partial_states(&Kernels_, &batch, &states) != 0) throw TDriverError("partial states failed");
        AdoptOutput(states);
        Pending_.push_back(states);
        return;
    }
    HashedRows_ += (uint64_t)batch.RowCount;
    Table_->Consume(batch);
    const int64_t groups = Table_->Groups();
    // Every group this lane produced counts, emitted or held.
    const double produced = (double)(EmittedGroups_ + (uint64_t)groups);
    if (PassThrough_ && HashedRows_ >= Policy_.SampleRows && (double)HashedRows_ < (double)Policy_.MinReduction * produced) {
        PassingThrough_ = true;
        Flush();
        return;
    }
    if (groups >= Policy_.MaxGroups) {
        EmittedGroups_ += (uint64_t)groups;
        Flush();
    }
}

bool TPartialAggregateDriver::Next(TRowSet& out)
{
    if (Pending_.empty()) return false;
    out = Pending_.front();
    Pending_.pop_front();
    return true;
}

bool TPartialAggregateDriver::Finish(TRowSet& out)
{
    if (PassingThrough_) return false;
    Table_->Emit(out);
    return true;
}

// ---------------------------------------------------------------------------
// Row join
// ---------------------------------------------------------------------------

TJoinDriver::TJoinDriver(const QdbRowJoinKernels& kernels, int64_t outputRows)
    : Kernels_(kernels)
    , OutputRows_(outputRows > 0 ? outputRows : 16384)
{
    if (!Op(QDB_JR_OP_INIT, nullptr)) throw TDriverError("row join initialization failed");
}

TJoinDriver::~TJoinDriver()
{
    for (TRowSet& rs : Outputs_) ReleaseOutput(rs);
    qdb_gpu_join_destroy(Arena_.data(), Table_.data(), Scratch_.data(), &Pairs_);
}

bool TJoinDriver::Op(int64_t op, const TRowSet* batch)
{
    TRowSet none{};
    TRowSet view = batch ? *batch : none;
    return Kernels_.Dispatch(Arena_.data(), Table_.data(), Scratch_.data(), &Pairs_, &view, nullptr, op);
}

void TJoinDriver::AddBuild(const TRowSet& batch)
{
    if (Sealed_) throw std::logic_error("row join append after build");
    if (!Op(QDB_JR_OP_APPEND, &batch)) throw TDriverError("row join append failed");
}

void TJoinDriver::Seal()
{
    if (Sealed_) return;
    if (!Op(QDB_JR_OP_BUILD, nullptr)) throw TDriverError("row join table build failed");
    Sealed_ = true;
}

void TJoinDriver::SealMany(TJoinDriver* const* drivers, int64_t n)
{
    std::vector<void*> arenas, tables;
    std::vector<TJoinDriver*> open;
    for (int64_t i = 0; i < n; ++i) {
        TJoinDriver* d = drivers[i];
        if (d->Kernels_.Slot != drivers[0]->Kernels_.Slot) throw std::invalid_argument("batched drivers of different joins");
        if (d->Sealed_) continue;
        arenas.push_back(d->Arena_.data());
        tables.push_back(d->Table_.data());
        open.push_back(d);
    }
    if (open.empty()) return;
    if (qdb_gpu_join_build_batched(&open[0]->Kernels_, arenas.data(), tables.data(), (int64_t)open.size()) != 0) {
        throw TDriverError("batched row join table build failed");
    }
    for (TJoinDriver* d : open) d->Sealed_ = true;
}

void TJoinDriver::ProbeMany(TJoinDriver* const* drivers, const TRowSet* const* batches, int64_t n)
{
    if (n <= 0) return;
    std::vector<void*> arenas((size_t)n), pairs((size_t)n);
    for (int64_t i = 0; i < n; ++i) {
        TJoinDriver* d = drivers[i];
        if (d->Kernels_.Slot != drivers[0]->Kernels_.Slot) throw std::invalid_argument("batched drivers of different joins");
        if (!d->Sealed_) throw std::logic_error("row join probe before build");
        arenas[(size_t)i] = d->Arena_.data();
        pairs[(size_t)i] = &d->Pairs_;
    }
    if (qdb_gpu_join_probe_batched(&drivers[0]->Kernels_, arenas.data(), batches, pairs.data(), n) != 0) {
        throw TDriverError("batched row join probe failed");
    }
}

int64_t TJoinDriver::Probe(const TRowSet& batch, bool materialize)
{
    if (!Sealed_) throw std::logic_error("row join probe before build");
    if (!Op(QDB_JR_OP_PROBE, &batch)) throw TDriverError("row join probe failed");
    const int64_t pairs = Pairs_.Count;
    if (!materialize) return pairs;
    // The pairs name rows of this batch: gather them while it is alive, in
    // output batches of OutputRows_ rows, then drain the pair buffer.
    TRowSet stream = batch;
    for (int64_t at = 0; at < pairs;) {
        TRowSet out{};
        const int64_t made = Kernels_.Materialize(&Pairs_, &stream, at, OutputRows_, &out);
        if (made <= 0) {
            ReleaseOutput(out);
            throw TDriverError("row join materialize failed");
        }
        AdoptOutput(out);
        Outputs_.push_back(out);
        at += made;
    }
    Pairs_.Count = 0;
    return pairs;
}

bool TJoinDriver::Next(TRowSet& out)
{
    if (Outputs_.empty()) return false;
    out = Outputs_.front();
    Outputs_.pop_front();
    return true;
}

}  // namespace NQdbGpu

// ---------------------------------------------------------------------------
// C exports. Errors: -1 / NULL, text in qdb_gpu_exec_last_error.
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

using NQdbGpu::TAggregateDriver;
using NQdbGpu::TJoinDriver;

}  // namespace

extern "C" {

const char* qdb_gpu_exec_last_error(void)
{
    return g_exec_error.c_str();
}

void* qdb_gpu_agg_processor_new(const QdbAggregateKernels* kernels, int64_t initial_capacity, int row_layout,
    int64_t initial_groups)
{
    TAggregateDriver* d = nullptr;
    Guard([&] {
        NQdbGpu::TAggregatePlan plan;
        plan.RowLayout = row_layout != 0;
        plan.InitialCapacity = initial_capacity;
        plan.InitialGroups = initial_groups;
        d = new TAggregateDriver(*kernels, plan);
    });
    return d;
}

int qdb_gpu_agg_processor_add(void* p, const TRowSet* batch)
{
    return Guard([&] { static_cast<TAggregateDriver*>(p)->Consume(*batch); });
}

int qdb_gpu_agg_processor_add_batched(void* const* procs, const TRowSet* const* batches, int64_t n)
{
    return Guard([&] { TAggregateDriver::ConsumeMany(reinterpret_cast<TAggregateDriver* const*>(procs), batches, n); });
}

int64_t qdb_gpu_agg_processor_finish(void* p, TRowSet* out)
{
    int64_t rows = -1;
    if (Guard([&] { rows = static_cast<TAggregateDriver*>(p)->Emit(*out); }) != 0) return -1;
    return rows;
}

int qdb_gpu_agg_processor_finish_batched(void* const* procs, TRowSet* outs, int64_t n)
{
    return Guard([&] { TAggregateDriver::EmitMany(reinterpret_cast<TAggregateDriver* const*>(procs), outs, n); });
}

int64_t qdb_gpu_agg_processor_groups(void* p)
{
    return static_cast<TAggregateDriver*>(p)->Groups();
}

void qdb_gpu_agg_processor_free(void* p)
{
    delete static_cast<TAggregateDriver*>(p);
}

void* qdb_gpu_join_processor_new(const QdbRowJoinKernels* kernels, int build_is_left, int64_t output_batch_rows)
{
    // The output column order is the compiled spec's (ProbeIsLeft); the
    // build side is the spec's build schema whichever side it is.
    (void)build_is_left;
    TJoinDriver* d = nullptr;
    Guard([&] { d = new TJoinDriver(*kernels, output_batch_rows); });
    return d;
}

int qdb_gpu_join_processor_append(void* p, const TRowSet* batch)
{
    return Guard([&] { static_cast<TJoinDriver*>(p)->AddBuild(*batch); });
}

int qdb_gpu_join_processor_build(void* p)
{
    return Guard([&] { static_cast<TJoinDriver*>(p)->Seal(); });
}

int64_t qdb_gpu_join_processor_probe(void* p, const TRowSet* batch)
{
    int64_t pairs = -1;
    if (Guard([&] { pairs = static_cast<TJoinDriver*>(p)->Probe(*batch, true); }) != 0) return -1;
    return pairs;
}

int64_t qdb_gpu_join_processor_probe_only(void* p, const TRowSet* batch)
{
    int64_t pairs = -1;
    if (Guard([&] { pairs = static_cast<TJoinDriver*>(p)->Probe(*batch, false); }) != 0) return -1;
    return pairs;
}

int qdb_gpu_join_processor_build_batched(void* const* procs, int64_t n)
{
    return Guard([&] { TJoinDriver::SealMany(reinterpret_cast<TJoinDriver* const*>(procs), n); });
}

int qdb_gpu_join_processor_probe_batched(void* const* procs, const TRowSet* const* batches, int64_t n)
{
    return Guard([&] { TJoinDriver::ProbeMany(reinterpret_cast<TJoinDriver* const*>(procs), batches, n); });
}

const QdbPairBuffer* qdb_gpu_join_processor_pairs(void* p)
{
    return &static_cast<TJoinDriver*>(p)->Pairs();
}

void* qdb_gpu_join_processor_arena(void* p)
{
    return static_cast<TJoinDriver*>(p)->Arena();
}

int qdb_gpu_join_processor_next(void* p, TRowSet* out)
{
    return static_cast<TJoinDriver*>(p)->Next(*out) ? 1 : 0;
}

void qdb_gpu_join_processor_free(void* p)
{
    delete static_cast<TJoinDriver*>(p);
}

void* qdb_gpu_partial_agg_new(const QdbAggregateKernels* kernels, int64_t initial_capacity, uint64_t sample_rows,
    uint64_t min_reduction, int64_t max_groups, int pass_through)
{
    NQdbGpu::TPartialAggregateDriver* d = nullptr;
    Guard([&] {
        NQdbGpu::TAggregatePlan plan;
        plan.InitialCapacity = initial_capacity;
        NQdbGpu::TPartialPolicy policy;
        policy.SampleRows = sample_rows;
        policy.MinReduction = min_reduction;
        policy.MaxGroups = max_groups;
        d = new NQdbGpu::TPartialAggregateDriver(*kernels, plan, policy, pass_through != 0);
    });
    return d;
}

int qdb_gpu_partial_agg_add(void* p, const TRowSet* batch)
{
    return Guard([&] { static_cast<NQdbGpu::TPartialAggregateDriver*>(p)->Consume(*batch); });
}

int qdb_gpu_partial_agg_next(void* p, TRowSet* out)
{
    return static_cast<NQdbGpu::TPartialAggregateDriver*>(p)->Next(*out) ? 1 : 0;
}

int64_t qdb_gpu_partial_agg_finish(void* p, TRowSet* out)
{
    int64_t rows = -1;
    if (Guard([&] {
            memset(out, 0, sizeof(*out));
            rows = static_cast<NQdbGpu::TPartialAggregateDriver*>(p)->Finish(*out) ? out->RowCount : 0;
        }) != 0) {
        return -1;
    }
    return rows;
}

int qdb_gpu_partial_agg_passing_through(void* p)
{
    return static_cast<NQdbGpu::TPartialAggregateDriver*>(p)->PassingThrough() ? 1 : 0;
}

void qdb_gpu_partial_agg_free(void* p)
{
    auto* d = static_cast<NQdbGpu::TPartialAggregateDriver*>(p);
    TRowSet rs{};
    while (d->Next(rs)) qdb_gpu_rowset_release(&rs);
    delete d;
}

void qdb_gpu_rowset_release(TRowSet* rs)
{
    if (!rs) return;
    if (rs->Destroy) rs->Destroy(rs);
    else qdb_gpu_destroy_kernel_rowset(rs);
}

}  // extern "C"
