// Filter, project and group-hash kernels.
//
// Filter: GenFilterKernelAst (qdb/kernel/gen.cpp:911-1091), called as
// filter(rs, arena, regexes, input_sel) (qdb/kernel/compiler.cpp:2252-2260,
// qdb/exec/filter_exec.cpp:27-70): rs->Selection[i] = the predicate IS TRUE
// for the rows input_sel selects (all rows when NULL), 0 for the others.
// Project: GenProjectKernelAst (gen.cpp:1373-1613), project(in, out, arena,
// regexes, input_sel) (compiler.cpp:2329-2336): out[k] data buffers,
// out[n + k] validity of nullable outputs; every row is evaluated when
// input_sel is NULL, else the selected rows.
// Group hash: CompileAggregateHash (compiler.cpp:3057-3125), rh_hash of the
// key columns of every selected row (HashKeyValue, gen.cpp:233-379).
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>
#include <vector>

#include "device.h"
#include "jit.h"
#include "runtime.h"
#include "slots.h"

namespace qdbgpu {

// The filter kernel's source for one predicate (generated beside the
// aggregate kernels, aggregate.hip).
std::string filter_jit_source(const DevProgram& prog, const QdbValueType* slot_type);
std::string project_jit_source(const DevProgram* progs, int nout, const QdbValueType* slot_type, const QdbValueType* out_types);

namespace {

constexpr int kBlock = 256;

struct ExprPlan {
    QdbColumnSchema input;
    int nslots = 0;
    int slot_col[kMaxDevCols];
    QdbValueType slot_type[kMaxDevCols];
    int nprogs = 0;
    DevProgram* d_progs = nullptr;
    std::vector<DevProgram> h_progs;
    QdbValueType out_types[QDB_PROJECT_MAX_OUTPUTS];
    // Digits each output may hold (DECIMAL(p <= 18); 0 = no check) and the
    // device word the kernel flags row errors in.
    int32_t out_prec[QDB_PROJECT_MAX_OUTPUTS] = {};
    bool checks = false;
    unsigned int* d_err = nullptr;
    // The kernel generated for a filter's predicate or a projection's
    // outputs (made at its first call; null: the generic interpreter kernel
    // runs).
    std::once_flag jit_once;
    hipFunction_t jit_filter = nullptr;
    hipFunction_t jit_project = nullptr;
    ~ExprPlan()
    {
        dfree(d_progs);
        dfree(d_err);
    }
};

// Remaps the programs' input columns to device slots and uploads them.
bool compile_programs(const QdbColumnSchema& input, const QdbExprProgram* progs, int n, ExprPlan* P, bool on_device = true)
{
    P->input = input;
    int col_slot[64];
    for (int c = 0; c < 64; ++c) col_slot[c] = -1;
    std::vector<QdbExprProgram> out((size_t)n);
    for (int p = 0; p < n; ++p) {
        out[(size_t)p] = progs[p];
        QdbExprProgram& q = out[(size_t)p];
        if (q.NumInsns < 0 || q.NumInsns > QDB_EXPR_MAX_INSNS) return fail("program has %d instructions", q.NumInsns);
        if (q.Result < 0 || q.Result >= QDB_EXPR_MAX_REGS) return fail("program result register out of range");
        for (int k = 0; k < q.NumInsns; ++k) {
            QdbExprInsn& in = q.Insns[k];
            if (in.Dst >= QDB_EXPR_MAX_REGS || in.A >= QDB_EXPR_MAX_REGS || in.B >= QDB_EXPR_MAX_REGS) {
                return fail("program register out of range");
            }
            if (in.Op == QDB_EXPR_DIV && in.Class != QDB_REG_F64) return fail("DIV is an f64 division (register class F64)");
            if (in.Op != QDB_EXPR_COLUMN) continue;
            if (in.Column < 0 || in.Column >= input.NumColumns) return fail("program column %d out of range", in.Column);
            const QdbValueType& t = input.Types[in.Column];
            if (t.Kind == QDB_KIND_STRING) return fail("string columns are not supported on this device");
            if (col_slot[in.Column] < 0) {
                if (P->nslots >= kMaxDevCols) return fail("too many columns in one program (max %d)", kMaxDevCols);
                col_slot[in.Column] = P->nslots;
                P->slot_type[P->nslots] = t;
                P->slot_col[P->nslots++] = in.Column;
            }
            in.Column = col_slot[in.Column];
        }
    }
    P->nprogs = n;
    if (n > 0) {
        std::vector<DevProgram> dev((size_t)n);
        for (int p = 0; p < n; ++p) dev[(size_t)p] = to_dev_program(out[(size_t)p]);
        P->h_progs = dev;
        if (!on_device) return true;
        P->d_progs = (DevProgram*)upload(dev.data(), dev.size() * sizeof(DevProgram));
        if (!P->d_progs) return false;
    }
    return true;
}

bool bind_cols(const ExprPlan* P, const TRowSet* rs, DevCols* cols)
{
    memset(cols, 0, sizeof(*cols));
    for (int s = 0; s < P->nslots; ++s) {
        const int c = P->slot_col[s];
        if (c >= rs->ColumnCount) return fail("batch has %lld columns, the program reads column %d", (long long)rs->ColumnCount, c);
        const TColumn& col = rs->Columns[c];
        const QdbValueType& t = P->input.Types[c];
        DevCol& d = cols->c[s];
        d.data = col.Data;
        d.mask = col.Mask;
        d.mask_off = col.MaskBitOffset;
        d.bit_off = col.DataBitOffset;
        d.kind = t.Kind;
        d.width = t.Width;
        d.nullable = t.Nullable;
    }
    return true;
}

__global__ void __launch_bounds__(kBlock) k_filter(DevCols cols, const DevProgram* prog, int64_t n,
    const uint8_t* in_sel, uint8_t* out_sel)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        uint8_t keep = 0;
        if (!in_sel || in_sel[i]) {
            DVal v = eval_program<true>(prog, cols, i);
            keep = (v.valid && v.i != 0) ? 1 : 0;
        }
        out_sel[i] = keep;
    }
}

__global__ void __launch_bounds__(kBlock) k_project(DevCols cols, const DevProgram* progs, int nout, ProjectOut O,
    int64_t n, const uint8_t* in_sel)
{
    const int lane = lane_id();
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // Waves cover 64 consecutive rows at multiples of 64 so that bitmap words
    // are written whole by one lane.
    for (int64_t base = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) - lane; base < n; base += stride) {
        const int64_t i = base + lane;
        const bool in_range = i < n;
        const bool eval = in_range && (!in_sel || in_sel[i]);
        const uint64_t evaluated = __ballot(eval);
#pragma unroll 1
        for (int k = 0; k < nout; ++k) {
            DVal v;
            v.valid = 0;
            v.i = 0;
            v.f = 0.0;
            v.cls = 0;
            if (eval) v = eval_program<true>(progs + k, cols, i);
            if (O.kind[k] == QDB_KIND_BOOL) {
                const uint64_t bits = __ballot(eval && v.valid && v.i != 0);
                if (lane == 0) store_bits((uint8_t*)O.data[k], base, n, bits, evaluated);
            } else if (eval) {
                char* d = O.data[k];
                if (O.kind[k] == QDB_KIND_FLOAT) {
                    double f = v.cls == QDB_REG_F64 ? v.f : (double)v.i;
                    if (!v.valid) f = 0.0;
                    if (O.width[k] == 4) ((float*)d)[i] = (float)f;
                    else ((double*)d)[i] = f;
                } else {
                    u128 x = v.valid ? (u128)v.i : 0;
                    if (O.prec[k] && v.valid && narrow_overflows(v.i, O.prec[k])) {
                        x = 0;
                        atomicOr(O.err, 1u << k);
                    }
                    switch (O.width[k]) {
                    case 1: ((uint8_t*)d)[i] = (uint8_t)x; break;
                    case 2: ((uint16_t*)d)[i] = (uint16_t)x; break;
                    case 4: ((uint32_t*)d)[i] = (uint32_t)x; break;
                    case 8: ((uint64_t*)d)[i] = (uint64_t)x; break;
                    default: {
                        uint64_t* p = (uint64_t*)(d + 16 * i);
                        p[0] = (uint64_t)x;
                        p[1] = (uint64_t)(x >> 64);
                    }
                    }
                }
            }
            if (O.mask[k]) {
                const uint64_t bits = __ballot(eval && v.valid);
                if (lane == 0) store_bits(O.mask[k], base, n, bits, evaluated);
            }
        }
    }
}

struct HashPlan {
    QdbColumnSchema input;
    int nkeys = 0;
};

struct HashCols {
    DevCol c[QDB_AGG_MAX_KEYS];
};

__global__ void __launch_bounds__(kBlock) k_group_hash(HashCols cols, int nkeys, int64_t n, const uint8_t* sel,
    uint64_t* hashes)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        // A rejected row gets no hash: its slot keeps what the caller left
        // there (qdb/kernel/join_gen.cpp:1557-1565).
        if (sel && !sel[i]) continue;
        hashes[i] = rh_key_hash(cols.c, nkeys, i);
    }
}

SlotTable<ExprPlan>& filter_slots()
{
    static SlotTable<ExprPlan>* t = new SlotTable<ExprPlan>();
    return *t;
}

SlotTable<ExprPlan>& project_slots()
{
    static SlotTable<ExprPlan>* t = new SlotTable<ExprPlan>();
    return *t;
}

SlotTable<HashPlan>& hash_slots()
{
    static SlotTable<HashPlan>* t = new SlotTable<HashPlan>();
    return *t;
}

void filter_impl(ExprPlan* P, TRowSet* rs, const uint8_t* input_sel)
{
    clear_error();
    if (!P || !rs || rs->RowCount <= 0) return;
    if (!rs->Selection) {
        fail("filter: the rowset needs a device Selection buffer for the output");
        return;
    }
    DevCols cols;
    if (!bind_cols(P, rs, &cols)) return;
    std::call_once(P->jit_once, [P]() {
        if (!jit_enabled() || P->nprogs != 1) return;
        std::string err;
        P->jit_filter = jit_function(filter_jit_source(P->h_progs[0], P->slot_type), "qdb_filter", &err);
    });
    {
        TimedRegion timer("filter");
        const int grid = grid_for(rs->RowCount, kBlock, 8);
        if (P->jit_filter) {
            int64_t n = rs->RowCount;
            const uint8_t* in_sel = input_sel;
            uint8_t* out_sel = rs->Selection;
            void* args[] = {&cols, &n, &in_sel, &out_sel};
            if (!hip_ok(hipModuleLaunchKernel(P->jit_filter, (unsigned)grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr),
                    "filter launch")) {
                return;
            }
        } else {
            hipLaunchKernelGGL(k_filter, dim3(grid), dim3(kBlock), 0, stream(), cols, P->d_progs, rs->RowCount, input_sel,
                rs->Selection);
        }
    }
    // Stream-ordered: the selection is ready for the library's next call
    // on its stream, and for the host after a copy or qdb_gpu_synchronize.
    (void)hip_ok(hipGetLastError(), "filter launch");
}

thread_local int64_t g_row_error = 0;
QdbRowErrorFn g_row_error_handler = nullptr;

// A row of the last project call failed: its precision is kept for
// qdb_gpu_take_row_error, the error text set, and the registered handler
// called (the reference's qdb_decimal64_overflow raises, or returns while
// the caller defers row errors).
void report_row_error(int precision)
{
    g_row_error = precision;
    fail("value does not fit DECIMAL precision %d", precision);
    if (g_row_error_handler) g_row_error_handler(precision);
}

void project_impl(ExprPlan* P, TRowSet* in, void** out, const uint8_t* input_sel)
{
    clear_error();
    g_row_error = 0;
    if (!P || !in || in->RowCount <= 0) return;
    DevCols cols;
    if (!bind_cols(P, in, &cols)) return;
    ProjectOut O;
    memset(&O, 0, sizeof(O));
    for (int k = 0; k < P->nprogs; ++k) {
        O.data[k] = (char*)out[k];
        O.mask[k] = P->out_types[k].Nullable ? (uint8_t*)out[P->nprogs + k] : nullptr;
        O.kind[k] = P->out_types[k].Kind;
        O.width[k] = P->out_types[k].Width;
        O.prec[k] = P->out_prec[k];
    }
    if (P->checks) {
        if (!P->d_err) P->d_err = (unsigned int*)dalloc(256);
        if (!P->d_err || !zero_fill(P->d_err, 4)) return;
        O.err = P->d_err;
    }
    std::call_once(P->jit_once, [P]() {
        if (!jit_enabled() || P->nprogs < 1) return;
        std::string err;
        P->jit_project = jit_function(project_jit_source(P->h_progs.data(), P->nprogs, P->slot_type, P->out_types),
            "qdb_project", &err);
    });
    {
        TimedRegion timer("project");
        const int grid = grid_for(in->RowCount, kBlock, 8);
        if (P->jit_project) {
            int64_t n = in->RowCount;
            const uint8_t* in_sel = input_sel;
            void* args[] = {&cols, &O, &n, &in_sel};
            if (!hip_ok(hipModuleLaunchKernel(P->jit_project, (unsigned)grid, 1, 1, kBlock, 1, 1, 0, stream(), args, nullptr),
                    "project launch")) {
                return;
            }
        } else {
            hipLaunchKernelGGL(k_project, dim3(grid), dim3(kBlock), 0, stream(), cols, P->d_progs, P->nprogs, O, in->RowCount,
                input_sel);
        }
    }
    if (!hip_ok(hipGetLastError(), "project launch")) return;
    if (P->checks) {
        // A possible row error is known only after the kernel: one readback.
        unsigned int err = 0;
        if (!hip_ok(hipMemcpyAsync(&err, P->d_err, 4, hipMemcpyDeviceToHost, stream()), "row error readback")
            || !sync("project")) {
            return;
        }
        if (err) report_row_error(P->out_prec[__builtin_ctz(err)]);
    }
    // Otherwise stream-ordered, as the filter.
}

bool hash_impl(HashPlan* P, TRowSet* rs, uint64_t* hashes, const int64_t* key_cols)
{
    clear_error();
    if (!P || !rs) return false;
    if (rs->RowCount <= 0) return true;
    HashCols cols;
    memset(&cols, 0, sizeof(cols));
    for (int k = 0; k < P->nkeys; ++k) {
        const int64_t c = key_cols[k];
        if (c < 0 || c >= rs->ColumnCount || c >= P->input.NumColumns) return fail("hash key column %lld out of range", (long long)c);
        const TColumn& col = rs->Columns[c];
        const QdbValueType& t = P->input.Types[c];
        if (t.Kind == QDB_KIND_STRING) return fail("string keys are not supported on this device");
        DevCol& d = cols.c[k];
        d.data = col.Data;
        d.mask = col.Mask;
        d.mask_off = col.MaskBitOffset;
        d.bit_off = col.DataBitOffset;
        d.kind = t.Kind;
        d.width = t.Width;
        d.nullable = t.Nullable;
    }
    hipLaunchKernelGGL(k_group_hash, dim3(grid_for(rs->RowCount, kBlock, 8)), dim3(kBlock), 0, stream(), cols, P->nkeys,
        rs->RowCount, rs->Selection, hashes);
    return hip_ok(hipGetLastError(), "hash launch") && sync("hash");
}

template <int N>
struct FilterTramp {
    static void call(TRowSet* rs, void* arena, void** regexes, const uint8_t* input_sel)
    {
        (void)arena;
        (void)regexes;
        filter_impl(filter_slots().get(N), rs, input_sel);
    }
};

template <int N>
struct ProjectTramp {
    static void call(TRowSet* in, void** out, void* arena, void** regexes, const uint8_t* input_sel)
    {
        (void)arena;
        (void)regexes;
        project_impl(project_slots().get(N), in, out, input_sel);
    }
};

template <int N>
struct HashTramp {
    static bool call(TRowSet* rs, uint64_t* hashes, const int64_t* key_cols)
    {
        return hash_impl(hash_slots().get(N), rs, hashes, key_cols);
    }
};

const auto kFilterFns = make_tramp_table<QdbFilterFn, FilterTramp>(std::make_integer_sequence<int, kSlots>{});
const auto kProjectFns = make_tramp_table<QdbProjectFn, ProjectTramp>(std::make_integer_sequence<int, kSlots>{});
const auto kHashFns = make_tramp_table<QdbGroupHashFn, HashTramp>(std::make_integer_sequence<int, kSlots>{});

}  // namespace

void release_aggregate_slot(int slot);
void release_join_slot(int slot);

}  // namespace qdbgpu

using namespace qdbgpu;

extern "C" int qdb_gpu_compile_filter(const QdbFilterSpec* spec, QdbFilterFn* out)
{
    clear_error();
    if (!spec || !out) return fail("null argument"), -1;
    if (!ensure_init()) return -1;
    ExprPlan* P = new ExprPlan();
    if (!compile_programs(spec->Input, &spec->Predicate, 1, P)) {
        delete P;
        return -1;
    }
    const int slot = filter_slots().add(P);
    if (slot < 0) {
        delete P;
        return fail("all filter slots are in use"), -1;
    }
    *out = kFilterFns.fns[slot];
    return slot;
}

// The source of the kernel generated for a filter's predicate (no GPU
// needed); returns its length, writing at most cap - 1 bytes and a NUL to
// out; -1 on error.
extern "C" int64_t qdb_gpu_filter_kernel_source(const QdbFilterSpec* spec, char* out, int64_t cap)
{
    clear_error();
    if (!spec) return fail("null argument"), -1;
    ExprPlan P;
    if (!compile_programs(spec->Input, &spec->Predicate, 1, &P, false)) return -1;
    const std::string src = filter_jit_source(P.h_progs[0], P.slot_type);
    if (out && cap > 0) {
        const size_t n = src.size() < (size_t)cap - 1 ? src.size() : (size_t)cap - 1;
        memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return (int64_t)src.size();
}

namespace {

// Checks a project spec's output types and records them in the plan.
bool set_project_outputs(const QdbProjectSpec* spec, ExprPlan* P)
{
    if (spec->NumOutputs < 1 || spec->NumOutputs > QDB_PROJECT_MAX_OUTPUTS) return fail("bad output count %d", spec->NumOutputs);
    for (int k = 0; k < spec->NumOutputs; ++k) {
        P->out_types[k] = spec->OutputTypes[k];
        const QdbValueType& t = spec->OutputTypes[k];
        const bool ok_width = t.Kind == QDB_KIND_BOOL || t.Width == 1 || t.Width == 2 || t.Width == 4 || t.Width == 8 || t.Width == 16;
        if (!ok_width || t.Kind == QDB_KIND_STRING) return fail("project output %d: unsupported type", k);
        if (t.Kind == QDB_KIND_DECIMAL && t.Width == 8) {
            const int p = t.Precision ? t.Precision : 18;
            if (p > 18) return fail("project output %d: DECIMAL(%d) does not fit 8 bytes", k, p);
            P->out_prec[k] = p;
            P->checks = true;
        }
    }
    return true;
}

}  // namespace

// The source of the kernel generated for a projection (no GPU needed), as
// qdb_gpu_filter_kernel_source.
extern "C" int64_t qdb_gpu_project_kernel_source(const QdbProjectSpec* spec, char* out, int64_t cap)
{
    clear_error();
    if (!spec) return fail("null argument"), -1;
    ExprPlan P;
    if (!set_project_outputs(spec, &P)) return -1;
    if (!compile_programs(spec->Input, spec->Outputs, spec->NumOutputs, &P, false)) return -1;
    const std::string src = project_jit_source(P.h_progs.data(), P.nprogs, P.slot_type, P.out_types);
    if (out && cap > 0) {
        const size_t n = src.size() < (size_t)cap - 1 ? src.size() : (size_t)cap - 1;
        memcpy(out, src.data(), n);
        out[n] = '\0';
    }
    return (int64_t)src.size();
}

extern "C" int qdb_gpu_compile_project(const QdbProjectSpec* spec, QdbProjectFn* out)
{
    clear_error();
    if (!spec || !out) return fail("null argument"), -1;
    if (spec->NumOutputs < 1 || spec->NumOutputs > QDB_PROJECT_MAX_OUTPUTS) return fail("bad output count %d", spec->NumOutputs), -1;
    if (!ensure_init()) return -1;
    ExprPlan* P = new ExprPlan();
    if (!set_project_outputs(spec, P)) {
        delete P;
        return -1;
    }
    if (!compile_programs(spec->Input, spec->Outputs, spec->NumOutputs, P)) {
        delete P;
        return -1;
    }
    const int slot = project_slots().add(P);
    if (slot < 0) {
        delete P;
        return fail("all project slots are in use"), -1;
    }
    *out = kProjectFns.fns[slot];
    return slot;
}

extern "C" void qdb_gpu_set_row_error_handler(QdbRowErrorFn fn)
{
    g_row_error_handler = fn;
}

extern "C" int64_t qdb_gpu_take_row_error(void)
{
    const int64_t p = g_row_error;
    g_row_error = 0;
    return p;
}

extern "C" int qdb_gpu_compile_group_hash(const QdbColumnSchema* input, int32_t num_keys, QdbGroupHashFn* out)
{
    clear_error();
    if (!input || !out) return fail("null argument"), -1;
    if (num_keys < 1 || num_keys > QDB_AGG_MAX_KEYS) return fail("bad key count %d", num_keys), -1;
    if (!ensure_init()) return -1;
    HashPlan* P = new HashPlan();
    P->input = *input;
    P->nkeys = num_keys;
    const int slot = hash_slots().add(P);
    if (slot < 0) {
        delete P;
        return fail("all hash slots are in use"), -1;
    }
    *out = kHashFns.fns[slot];
    return slot;
}

extern "C" void qdb_gpu_release_slot(int32_t kind, int32_t slot)
{
    switch (kind) {
    case QDB_SLOT_FILTER: filter_slots().release(slot); break;
    case QDB_SLOT_PROJECT: project_slots().release(slot); break;
    case QDB_SLOT_AGGREGATE: release_aggregate_slot(slot); break;
    case QDB_SLOT_JOIN: release_join_slot(slot); break;
    case QDB_SLOT_HASH: hash_slots().release(slot); break;
    default: break;
    }
}
