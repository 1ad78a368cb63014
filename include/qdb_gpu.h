/*
 * qdb_gpu.h - C-ABI of the MI355X (gfx950) hash-aggregate / hash-join path.
 *
 * The reference engine calls its operator kernels through plain function
 * pointers held in a kernel slot (qdb/kernel/generated.h:82-128): arguments
 * are pointers or integers of at most 64 bits, results void/int/bool. This
 * library fills such slots with entry points that run hand-written HIP
 * kernels. Every signature below keeps the reference's argument meaning,
 * op-codes, state-blob sizes and error convention:
 *
 *   agg_dispatch      int64_t (void* ht, TRowSet* batch, int64_t arg, int64_t op)
 *                     qdb/kernel/compiler.h:208-212, compiler.cpp:2776-2782
 *   agg_finish_rowset int64_t (void* ht, TRowSet* out)
 *                     qdb/kernel/compiler.h:232-233
 *   jr_dispatch       bool (void* arena, void* table, void* scratch, void* pairs,
 *                           TRowSet* batch, const int64_t* key_columns, int64_t op)
 *                     qdb/kernel/compiler.h:385-393, compiler.cpp:3242-3265
 *   jr_materialize    int64_t (void* pairs, TRowSet* stream, int64_t start,
 *                              int64_t limit, TRowSet* out)
 *                     qdb/kernel/compiler.cpp:3244,3266-3274
 *   filter <kernel>   void (TRowSet* rs, void* arena, void** regexes,
 *                           const uint8_t* input_sel)
 *                     qdb/kernel/compiler.cpp:2252-2260
 *   <project>         void (TRowSet* in, void** out, void* arena, void** regexes,
 *                           const uint8_t* input_sel)
 *                     qdb/kernel/compiler.cpp:2329-2336
 *   group hash        bool (TRowSet* rs, uint64_t* hashes, const int64_t* key_cols)
 *                     qdb/kernel/compiler.cpp:3057-3125
 *
 * The reference generates one native function per query; here a "compile"
 * call binds a query spec to a slot of a fixed pool of trampolines and
 * returns plain function pointers with exactly those signatures, which is
 * what FinalizeKernels stores into TKernelSlot::Fns.
 *
 * Column buffers handed to these entry points live in device memory (HBM):
 * qdb_gpu_stage_rowset copies host batches in once through pinned buffers.
 * Every entry point is synchronous with respect to its caller, as the
 * reference's CPU kernels are: when it returns, the batch may be released.
 */
#ifndef QDB_GPU_H
#define QDB_GPU_H

#ifndef __HIPCC_RTC__
#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#else
/* Kernels compiled at run time (hipRTC) get the fixed-width types from the
 * runtime compiler's own header. */
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int16_t int16_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
typedef __hip_internal::uint8_t uint8_t;
typedef __hip_internal::uint16_t uint16_t;
typedef __hip_internal::uint32_t uint32_t;
typedef __hip_internal::uint64_t uint64_t;
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Column-batch ABI, byte-compatible with qdb/io/io.h:17-39.                 */
/* ------------------------------------------------------------------------ */

typedef struct TColumn {
    char* Data;
    int32_t DataBitOffset;   /* bit offset for bit-packed (bool) columns */
    const uint8_t* Mask;     /* Arrow LSB validity bitmap; NULL = all valid */
    int32_t MaskBitOffset;
    void* Offsets;           /* variable-length offsets; NULL for fixed width */
    uint8_t OffsetWidth;     /* 4 or 8 for strings, 0 otherwise */
} TColumn;

typedef struct TRowSet {
    TColumn* Columns;
    int64_t ColumnCount;
    int64_t RowCount;
    uint8_t* Selection;      /* NULL = all rows; Selection[i] == 0 skips row i */
    uint64_t* Hash;          /* NULL, or rh_hash of the key of every selected row */
    void (*Destroy)(struct TRowSet*);
    void* Private;
    int64_t RefCount;
} TRowSet;

/* Value kinds, qdb/io/value_type.h:7-18. */
enum {
    QDB_KIND_SIGNED = 0,
    QDB_KIND_UNSIGNED = 1,
    QDB_KIND_FLOAT = 2,
    QDB_KIND_BOOL = 3,
    QDB_KIND_STRING = 4,
    QDB_KIND_DATE = 5,
    QDB_KIND_DECIMAL = 6
};

/* How a column's values are read, qdb/io/value_type.h:20-35. A DECIMAL is an
 * unscaled signed integer of Width 8 (precision <= 18) or 16 (i128),
 * qdb/plan/types/decimal.h:80-99. */
typedef struct QdbValueType {
    uint8_t Kind;
    int32_t Width;
    int32_t Scale;
    bool Nullable;
    /* DECIMAL precision (digits); 0 = the storage's widest: 18 for Width 8,
     * 38 for Width 16. A project output of DECIMAL(p <= 18) whose value has
     * more than p digits is a row error (qdb_decimal64_narrow_checked,
     * qdb/modules/qumirdb.oz:1236-1245). */
    uint8_t Precision;
} QdbValueType;

/* ------------------------------------------------------------------------ */
/* Op-codes and state-blob sizes (qdb/kernel/compiler.h:217-230,363-383,     */
/* 603-616).                                                                 */
/* ------------------------------------------------------------------------ */

enum {
    QDB_AGG_OP_INIT = 0,
    QDB_AGG_OP_UPDATE = 1,
    QDB_AGG_OP_INIT_ROWS = 3,
    QDB_AGG_OP_SWITCH_TO_ROWS_AT = 4,
    QDB_AGG_OP_EXPECT_ROW_GROUPS = 5
};

enum {
    QDB_JR_OP_INIT = 0,
    QDB_JR_OP_APPEND = 1,
    QDB_JR_OP_BUILD = 2,
    QDB_JR_OP_PROBE = 3,
    QDB_JR_OP_APPEND_NULL_KEYS = 5,
    QDB_JR_OP_EMIT_BY_MARK = 6
};

#define QDB_HASH_TABLE_SIZE 304
#define QDB_PAIR_BUFFER_SIZE 24
#define QDB_ROW_ARENA_SIZE 56
#define QDB_ROW_TABLE_SIZE 48
#define QDB_ROW_SCRATCH_SIZE 64
#define QDB_ROWSET_SIZE 64

/* The PairBuffer blob (qdb/exec/row_join_exec.h:95-100). Data holds 2 i64 per
 * pair, (left id, right id); here Data is device memory. A stream row id is
 * (-1 << 32) + row, as jr_push_pair writes it (qdb/kernel/join/join_rows.oz:
 * 117-125); the build side is the build row's ordinal over every appended
 * batch (the reference stores a row pointer there, which has no meaning on
 * another device). */
typedef struct QdbPairBuffer {
    int64_t Count;
    int64_t Capacity;
    int64_t* Data;
} QdbPairBuffer;

/* ------------------------------------------------------------------------ */
/* Expression programs: predicates (filter) and computed columns (project).  */
/* A program is a straight-line list of register instructions evaluated per */
/* row. Values are exact i128 (integers and unscaled decimals), f64, or a   */
/* three-valued bool; every register carries a validity flag (SQL NULL).    */
/* ------------------------------------------------------------------------ */

enum {
    QDB_EXPR_COLUMN = 1,   /* r[dst] = column[col] widened to i128 / f64 / bool   */
    QDB_EXPR_CONST_INT = 2,/* r[dst] = (imm_hi:imm_lo) as i128                    */
    QDB_EXPR_CONST_F64 = 3,/* r[dst] = f64 bits in imm_lo                         */
    QDB_EXPR_ADD = 4,      /* r[dst] = r[a] + r[b] (i128, wraps; or f64)           */
    QDB_EXPR_SUB = 5,
    QDB_EXPR_MUL = 6,
    QDB_EXPR_NEG = 7,
    QDB_EXPR_SCALE_UP = 8, /* r[dst] = r[a] * 10^imm_lo (decimal rescale)          */
    QDB_EXPR_CMP = 9,      /* r[dst] = r[a] <cmp imm_lo> r[b], three-valued        */
    QDB_EXPR_AND = 10,     /* SQL three-valued AND                                 */
    QDB_EXPR_OR = 11,
    QDB_EXPR_NOT = 12,
    QDB_EXPR_IS_NULL = 13, /* r[dst] = r[a] IS NULL (never NULL)                   */
    QDB_EXPR_TO_F64 = 14,  /* r[dst] = decimal r[a] at scale imm_lo as nearest f64,
                            * ties to even (qdb_decimal_to_f64, qdb/modules/
                            * qumirdb.oz:983-1046); an f64 operand is kept */
    QDB_EXPR_DIV = 15      /* r[dst] = r[a] / r[b], IEEE f64 division (class F64
                            * only): AVG = SUM / COUNT of the lowering,
                            * qdb/plan/sql_lowering.cpp:799-814 */
};

enum { QDB_CMP_EQ = 0, QDB_CMP_NE = 1, QDB_CMP_LT = 2, QDB_CMP_LE = 3, QDB_CMP_GT = 4, QDB_CMP_GE = 5 };

/* Register classes. */
enum { QDB_REG_INT = 0, QDB_REG_F64 = 1, QDB_REG_BOOL = 2 };

#define QDB_EXPR_MAX_INSNS 32
#define QDB_EXPR_MAX_REGS 8

typedef struct QdbExprInsn {
    uint8_t Op;
    uint8_t Dst;
    uint8_t A;
    uint8_t B;
    uint8_t Class;     /* register class of the result (QDB_REG_*) */
    uint8_t Pad[3];
    int32_t Column;    /* QDB_EXPR_COLUMN: input column index */
    int32_t Pad2;
    int64_t ImmLo;
    int64_t ImmHi;
} QdbExprInsn;

typedef struct QdbExprProgram {
    int32_t NumInsns;
    int32_t Result;    /* register holding the result */
    QdbExprInsn Insns[QDB_EXPR_MAX_INSNS];
} QdbExprProgram;

/* ------------------------------------------------------------------------ */
/* Filter kernel: GenFilterKernelAst (qdb/kernel/gen.cpp:911-1091).          */
/* sel[i] = (input_sel == NULL || input_sel[i]) ? (pred(row i) IS TRUE) : 0, */
/* written to rs->Selection (which must be a device buffer of RowCount      */
/* bytes). Rows rejected upstream are not evaluated.                         */
/* ------------------------------------------------------------------------ */

typedef struct QdbColumnSchema {
    int32_t NumColumns;
    QdbValueType Types[64];
} QdbColumnSchema;

typedef struct QdbFilterSpec {
    QdbColumnSchema Input;
    QdbExprProgram Predicate;   /* result register class QDB_REG_BOOL */
} QdbFilterSpec;

typedef void (*QdbFilterFn)(TRowSet* rs, void* arena, void** regexes, const uint8_t* input_sel);

/* ------------------------------------------------------------------------ */
/* Project kernel: GenProjectKernelAst (qdb/kernel/gen.cpp:1373-1613).       */
/* out[k] = data buffer of output k (RowCount * width bytes, device; a bool */
/* output is an LSB bitmap), out[NumOutputs + k] = its validity bitmap      */
/* (device) when the output is nullable (gen.cpp:1519-1535). Every row is   */
/* evaluated when input_sel is NULL; otherwise only the selected rows.      */
/* ------------------------------------------------------------------------ */

#define QDB_PROJECT_MAX_OUTPUTS 8

typedef struct QdbProjectSpec {
    QdbColumnSchema Input;
    int32_t NumOutputs;
    QdbValueType OutputTypes[QDB_PROJECT_MAX_OUTPUTS];
    QdbExprProgram Outputs[QDB_PROJECT_MAX_OUTPUTS];
} QdbProjectSpec;

typedef void (*QdbProjectFn)(TRowSet* in, void** out, void* arena, void** regexes, const uint8_t* input_sel);

/* ------------------------------------------------------------------------ */
/* Aggregate kernel: agg_dispatch / agg_finish_rowset                        */
/* (qdb/kernel/gen.cpp:2075-2299, 2836-3132).                                */
/* ------------------------------------------------------------------------ */

enum {
    QDB_AGG_COUNT_STAR = 0,
    QDB_AGG_COUNT = 1,
    QDB_AGG_SUM = 2,
    QDB_AGG_MIN = 3,
    QDB_AGG_MAX = 4
};

#define QDB_AGG_MAX_KEYS 4
#define QDB_AGG_MAX_AGGS 16
#define QDB_AGG_MAX_EMPTY_INPUT_GROUPS 4

typedef struct QdbAggregateFn {
    int32_t Func;       /* QDB_AGG_* */
    int32_t Column;     /* argument column; -1 for COUNT(*) */
} QdbAggregateFn;

typedef struct QdbAggregateSpec {
    QdbColumnSchema Input;
    int32_t NumKeys;
    int32_t KeyColumns[QDB_AGG_MAX_KEYS];
    int32_t NumAggs;
    QdbAggregateFn Aggs[QDB_AGG_MAX_AGGS];
    /* Groups that exist over an empty input (TOperatorKernelSpec::
     * EmptyInputGroups, qdb/kernel/spec.h:85-89; seeded by the generated
     * Init, qdb/kernel/gen.cpp:2218-2269): Init creates one group per entry
     * whose leading key column (a non-nullable integer) holds the entry and
     * whose other key columns are NULL (0 when not nullable), with every
     * reducer at its identity (COUNT 0, SUM/MIN/MAX NULL), so the table's
     * Size counts them from Init on. A SQL global aggregate groups by the
     * constant i64 column `__group__` = 1 and seeds {1}
     * (qdb/plan/build.cpp:2461-2524); its SUM/MIN/MAX results are nullable
     * (qdb/plan/ops/aggregate.cpp:93-97), as they are here whenever a group
     * is seeded. */
    int32_t NumEmptyInputGroups;
    int64_t EmptyInputGroups[QDB_AGG_MAX_EMPTY_INPUT_GROUPS];
    /* Output buffers of agg_finish_rowset: 0 = device memory, 1 = host. */
    int32_t OutputOnHost;
    /* Optional fused stages, each equal to the separate operator: a
     * predicate (filter kernel) and computed columns appended after the
     * input columns (project kernel). NumComputed = 0 disables project. */
    int32_t HasPredicate;
    QdbExprProgram Predicate;
    int32_t NumComputed;
    QdbValueType ComputedTypes[QDB_PROJECT_MAX_OUTPUTS];
    QdbExprProgram Computed[QDB_PROJECT_MAX_OUTPUTS];
} QdbAggregateSpec;

typedef int64_t (*QdbAggDispatchFn)(void* ht, TRowSet* batch, int64_t arg, int64_t op);
typedef int64_t (*QdbAggFinishFn)(void* ht, TRowSet* out);

/* TAggregateKernelMeta (qdb/exec/plan_data.h:173-205) plus the callables. */
typedef struct QdbAggregateKernels {
    QdbAggDispatchFn Dispatch;
    QdbAggFinishFn FinishRowSet;
    int64_t NumAggs;
    int64_t StateWords;
    int64_t RowBytes;       /* 0: no separate row layout on this device */
    int32_t Slot;
    /* Output schema of agg_finish_rowset: keys, then one column per agg. */
    int32_t NumOutputColumns;
    QdbValueType OutputTypes[QDB_AGG_MAX_KEYS + QDB_AGG_MAX_AGGS];
} QdbAggregateKernels;

/* ------------------------------------------------------------------------ */
/* Row join: jr_dispatch / jr_materialize (qdb/kernel/join_gen.cpp:2017-2375, */
/* qdb/kernel/join/join_rows.oz).                                            */
/* ------------------------------------------------------------------------ */

enum { QDB_JOIN_INNER = 0, QDB_JOIN_LEFT_OUTER = 1, QDB_JOIN_RIGHT_OUTER = 2 };

#define QDB_JOIN_MAX_KEYS 2
#define QDB_JOIN_MAX_FUSED 4

typedef struct QdbRowJoinSpec {
    QdbColumnSchema Build;
    QdbColumnSchema Probe;
    int32_t NumKeys;
    int32_t BuildKeyColumns[QDB_JOIN_MAX_KEYS];
    int32_t ProbeKeyColumns[QDB_JOIN_MAX_KEYS];
    int32_t JoinType;        /* QDB_JOIN_* */
    int32_t ProbeIsLeft;     /* output column order: probe columns first */
    /* Optional global aggregate over the matched pairs, fused into Probe
     * (COUNT(*) and SUM/MIN/MAX of a probe or build column), the same value
     * an aggregate over jr_materialize's output computes. Column >= 0 names a
     * probe column, Column <= -2 the build column (-Column - 2). */
    int32_t NumFused;
    QdbAggregateFn Fused[QDB_JOIN_MAX_FUSED];
    /* Output buffers of jr_materialize: 0 = device memory, 1 = host. */
    int32_t OutputOnHost;
    /* Build table: QDB_JOIN_TABLE_AUTO picks a direct table indexed by
     * key - min (a perfect hash) when the build has one integer key, distinct
     * keys and a key range of at most 8 x its rows (or 2^16); otherwise, or
     * with QDB_JOIN_TABLE_HASH, the open-addressing hash table. Results are
     * the same either way. */
    int32_t TableMode;
} QdbRowJoinSpec;

enum { QDB_JOIN_TABLE_AUTO = 0, QDB_JOIN_TABLE_HASH = 1 };

typedef bool (*QdbJoinDispatchFn)(void* arena, void* table, void* scratch, void* pairs,
    TRowSet* batch, const int64_t* key_columns, int64_t op);
typedef int64_t (*QdbJoinMaterializeFn)(void* pairs, TRowSet* stream, int64_t start,
    int64_t limit, TRowSet* out);

typedef struct QdbRowJoinKernels {
    QdbJoinDispatchFn Dispatch;
    QdbJoinMaterializeFn Materialize;
    int32_t Slot;
    int32_t NumOutputColumns;
    QdbValueType OutputTypes[128];
} QdbRowJoinKernels;

/* Fused global-aggregate results of the probes since Init, per fused agg:
 * Count = contributing rows, Value = i128 sum (lo, hi) or the f64 bits of
 * min/max/sum in Lo, Valid = 0 when the result is NULL. */
typedef struct QdbFusedResult {
    int64_t Lo;
    int64_t Hi;
    int64_t Count;
    int64_t Valid;
} QdbFusedResult;

/* ------------------------------------------------------------------------ */
/* Library entry points.                                                     */
/* ------------------------------------------------------------------------ */

/* Initialises the device runtime on device `device` (hipSetDevice, stream,
 * memory pool). Returns 0 on success. Every other call initialises lazily. */
int qdb_gpu_init(int device);
/* Text of the last error on this thread ("" when none). */
const char* qdb_gpu_last_error(void);
/* Library version string. */
const char* qdb_gpu_version(void);

/* Compile = bind a spec to a trampoline slot and return its entry points.
 * Return -1 on an unsupported spec (see qdb_gpu_last_error); on success the
 * filter, project and group-hash compilers return their slot index (>= 0,
 * for qdb_gpu_release_slot) and the aggregate and join compilers return 0
 * with the slot in the kernels' Slot field. */
int qdb_gpu_compile_filter(const QdbFilterSpec* spec, QdbFilterFn* out);
int qdb_gpu_compile_project(const QdbProjectSpec* spec, QdbProjectFn* out);
int qdb_gpu_compile_aggregate(const QdbAggregateSpec* spec, QdbAggregateKernels* out);
int qdb_gpu_compile_row_join(const QdbRowJoinSpec* spec, QdbRowJoinKernels* out);
/* The group-hash entry point: rh_hash of the key columns of every selected
 * row into `hashes` (device), qdb/kernel/compiler.cpp:3057-3125. */
typedef bool (*QdbGroupHashFn)(TRowSet* rs, uint64_t* hashes, const int64_t* key_cols);
int qdb_gpu_compile_group_hash(const QdbColumnSchema* input, int32_t num_keys, QdbGroupHashFn* out);
/* Row errors of the project kernel: an output of DECIMAL(p <= 18) whose
 * value has more than p digits (qdb_decimal64_narrow_checked,
 * qdb/modules/qumirdb.oz:1236-1245). As in the reference, the failing value
 * is stored as 0 and the row error goes to the runtime's overflow routine,
 * qdb_decimal64_overflow(p) (qdb/modules/qumirdb_runtime.cpp:662-667), which
 * raises or, while the caller defers row errors, returns 0
 * (qdb/exec/project_exec.cpp:71-81: every row evaluated deferred, then the
 * selected rows raising). Here the project entry point calls the registered
 * handler once per call, with the precision of the lowest-numbered failing
 * output, after its kernel; qdb_gpu_take_row_error returns that precision
 * (0: no row failed) and clears it, and the error text is
 * "value does not fit DECIMAL precision p". */
typedef int64_t (*QdbRowErrorFn)(int64_t precision);
void qdb_gpu_set_row_error_handler(QdbRowErrorFn fn);
int64_t qdb_gpu_take_row_error(void);
/* Frees the slot of a compiled kernel (its Slot field / the fn pointer). */
void qdb_gpu_release_slot(int32_t kind, int32_t slot);
enum { QDB_SLOT_FILTER = 0, QDB_SLOT_PROJECT = 1, QDB_SLOT_AGGREGATE = 2, QDB_SLOT_JOIN = 3, QDB_SLOT_HASH = 4 };

/* Fused join aggregate results (see QdbRowJoinSpec.NumFused); `arena` is the
 * join's RowArena blob. Returns the number of results written. */
int64_t qdb_gpu_join_fused_results(void* arena, QdbFusedResult* out, int64_t max_out);

/* The fused-aggregate probe that allocates nothing. It has no counterpart
 * in the reference; it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It does the work of jr_dispatch(Probe) followed by
 * qdb_gpu_join_fused_results for that one batch, with two differences: no
 * pair is written (the pair count goes to *pairs_out), and out[f] holds
 * fused aggregate f of this batch alone (not accumulated since Init, and not
 * added to what qdb_gpu_join_fused_results reports). `arena` is the RowArena
 * blob of a join after Build; max_out must be at least its NumFused. The
 * per-workgroup partials live in a static device array of 128 KiB per
 * device (room for 512 workgroups); a launch that would need more is refused
 * with an error. Every call on one device shares that array and the join's
 * own result words, so calls that overlap on one device (or on one join)
 * race. Calls no hipMalloc / hipFree. Returns 0, or -1 with the error text. */
int qdb_gpu_join_probe_fused_noalloc(void* arena, const TRowSet* batch, QdbFusedResult* out, int64_t max_out,
    int64_t* pairs_out);

/* Frees everything behind a join's arena/table/scratch/pairs blobs (the
 * reference frees them with its allocation domain). */
void qdb_gpu_join_destroy(void* arena, void* table, void* scratch, void* pairs);
/* Frees a table behind a HashTable blob without finalizing it. */
void qdb_gpu_aggregate_destroy(void* ht);

/* Device memory (from the library's caching pool) and copies. */
void* qdb_gpu_alloc(size_t bytes);
void qdb_gpu_free(void* ptr);
/* Returns the device blocks the library's allocator keeps for reuse to the
 * driver (otherwise it keeps freed blocks while 1/8 of the device memory
 * stays free). No
 * counterpart in the reference, whose kernels allocate from host heaps. */
void qdb_gpu_trim_cache(void);

/* Device pool counters, for diagnostics (no reference counterpart):
 * [driver allocations, driver frees, free-memory queries, cached bytes,
 * bytes in blocks handed out]; returns how many values were written. */
int64_t qdb_gpu_pool_stats(int64_t* out, int64_t n);
int qdb_gpu_memcpy_h2d(void* dst, const void* src, size_t bytes);
int qdb_gpu_memcpy_d2h(void* dst, const void* src, size_t bytes);
int qdb_gpu_memset(void* dst, int value, size_t bytes);
int qdb_gpu_synchronize(void);

/* The host heap of the reference's kernels: qdb_alloc / qdb_free
 * (qdb/modules/qumirdb_runtime.h:52-53), and TKernelHeap::Detach
 * (qdb/kernel/execution_context.h:39), which moves a block out of the task's
 * heap when an output is adopted (qdb/exec/kernel_rowset.h:30-38). Every
 * output rowset of agg_finish_rowset and jr_materialize is kernel-owned as
 * the reference's are: TColumns and host buffers come from `alloc`, Private
 * is the owners list (owners[0] = count, then the blocks; the list is an
 * `alloc` block too) and Destroy is NULL, for the host to adopt
 * (AdoptKernelOwnedRowSet) and release (DestroyKernelOwnedRowSet frees every
 * owner with `free`). An output left in device memory (OutputOnHost = 0) has
 * one more owner, a host block listing its device buffers, which only
 * qdb_gpu_destroy_kernel_rowset returns to the device pool: such outputs
 * feed device-side consumers, which release them with that call. Without a
 * registered heap, malloc / free are used and detach is a no-op. */
typedef void* (*QdbHeapAllocFn)(int64_t size);
typedef void (*QdbHeapFreeFn)(void* ptr);
typedef void (*QdbHeapDetachFn)(void* ptr);
void qdb_gpu_set_host_heap(QdbHeapAllocFn alloc, QdbHeapFreeFn free_fn, QdbHeapDetachFn detach);
/* DestroyKernelOwnedRowSet for this library's outputs (device buffers
 * included); callable as a TRowSet Destroy callback. */
void qdb_gpu_destroy_kernel_rowset(TRowSet* rs);
/* AdoptKernelOwnedRowSet (qdb/exec/kernel_rowset.h:30-38) for callers without
 * the reference's: detaches every owner (and the list) through the heap's
 * detach and installs qdb_gpu_destroy_kernel_rowset as Destroy. */
void qdb_gpu_adopt_kernel_rowset(TRowSet* rs);

/* Staging (qdb/io; decimal storage qdb/io/parquet/decimal_storage.cpp:220-247): copies a host TRowSet's fixed-width
 * columns, masks, selection and hashes into HBM through one pinned buffer
 * with hipMemcpyAsync (one synchronization), producing a device TRowSet
 * whose Destroy (qdb_gpu_destroy_kernel_rowset) frees it. `types` gives each
 * column's width. Returns 0 on success. */
int qdb_gpu_stage_rowset(const TRowSet* host, const QdbColumnSchema* types, TRowSet* device_out);
/* Copies a device TRowSet (as produced by agg_finish_rowset/jr_materialize)
 * to host memory (the host heap); the result's Destroy frees it. */
int qdb_gpu_download_rowset(const TRowSet* device, const QdbValueType* types, TRowSet* host_out);

/* Synthetic inputs generated in HBM from a seed (counter-based, identical to
 * the host generator in qdb_gen.h). Column buffers are caller-allocated. */
int qdb_gpu_gen_lineitem(uint64_t seed, int64_t row0, int64_t rows, int8_t* flag, int8_t* status,
    int32_t* shipdate, int64_t* quantity, int64_t* price, int64_t* discount, int64_t* tax);
int qdb_gpu_gen_join_build(uint64_t seed, int64_t row0, int64_t rows, int64_t* key, int32_t* segment);
int qdb_gpu_gen_join_probe(uint64_t seed, int64_t row0, int64_t rows, int64_t build_rows, int64_t* fk, int64_t* value);
int qdb_gpu_gen_zipf_groups(uint64_t seed, int64_t row0, int64_t rows, const double* zipf_cdf,
    int64_t distinct, int64_t* key, int64_t* vi, double* vf);

/* Runtime join filter (TRuntimeFilter, qdb/exec/runtime_filter.cpp:122-300;
 * its kernel rf_build / rf_publish / rf_probe, qdb/kernel/join/
 * runtime_filter.oz): the probe-side filter of a join, built from the
 * selected build rows' join-key hashes (rh_hash of the key struct, as the
 * group-hash entry point computes it). Defaults of the reference: max_keys
 * 16384 (TRuntimeFilter::DefaultMaxKeys), bloom_bytes 2 MiB (rounded down to
 * a power of two; below 64 bytes there is no Bloom storage). Up to max_keys
 * distinct hashes the published filter is an exact set (kind 1), beyond that
 * a blocked Bloom filter (kind 2), without Bloom storage it cannot answer
 * (kind 0: every selected row is kept). Hashes, selections and outputs are
 * device arrays of `count` entries; a NULL selection selects every row.
 *   qdb_gpu_rf_new      -> an opaque filter, or NULL
 *   qdb_gpu_rf_build    adds a build batch's selected hashes (0 or -1)
 *   qdb_gpu_rf_publish  publishes; returns the kind, or -1
 *   qdb_gpu_rf_probe    output[i] = 1 when row i is selected and its hash may
 *                       be in the build (rf_probe); returns the kept count
 *   qdb_gpu_rf_stats    [kind, exact keys, exact table slots, Bloom words,
 *                       Bloom bits set]; returns the count written */
void* qdb_gpu_rf_new(int64_t max_keys, int64_t bloom_bytes);
int qdb_gpu_rf_build(void* filter, const uint64_t* hashes, const uint8_t* selection, int64_t count);
int qdb_gpu_rf_publish(void* filter);
int64_t qdb_gpu_rf_probe(const void* filter, const uint64_t* hashes, const uint8_t* selection, uint8_t* output,
    int64_t count);
int64_t qdb_gpu_rf_stats(const void* filter, int64_t* out, int64_t n);
void qdb_gpu_rf_free(void* filter);

/* Hash shuffle routing (THashShuffleTask::OutputOf, qdb/scheduler/tasks/
 * runtime_adapter.cpp:998-1010): every selected row of `rs` goes to output
 * hash % outputs, or to routes[hash % nroutes] with a routing table (host
 * array; an entry outside [0, outputs) routes a partition nowhere, and a row
 * of such a partition is an error with the reference's text). The hash is
 * rs->Hash when the batch has one (computed upstream on the same key, the
 * reference's UseInputHash), else `hashes` (device, e.g. the group-hash entry
 * point's rh_hash). The routing is a stable device partition: row_index
 * (device, RowCount entries) lists the selected rows grouped by output, each
 * output's rows in input order, output o's at [offsets[o], offsets[o+1])
 * (offsets: host, outputs + 1 entries). At most 4096 outputs. 0 or -1. */
int qdb_gpu_shuffle_partition(const TRowSet* rs, const uint64_t* hashes, int32_t outputs, const int32_t* routes,
    int32_t nroutes, uint32_t* row_index, int64_t* offsets);
/* Materializes rows row_index[0..count) (device) of a device rowset as a new
 * device rowset: every column (bool data and validity as bitmaps), and the
 * hashes and selection when the batch has them; kernel-owned, Destroy set
 * (qdb_gpu_destroy_kernel_rowset). 0 or -1. */
int qdb_gpu_gather_rowset(const TRowSet* rs, const QdbColumnSchema* types, const uint32_t* row_index, int64_t count,
    TRowSet* out);

/* The single-row partial states of a batch for an aggregate: the projection
 * that the reference's adaptive partial aggregate passes rows through once
 * hashing stops paying (docs/arch/aggregation.md:376-389): the keys as they
 * are, 1 for COUNT(*), IS NOT NULL (0/1) for COUNT(x), the argument of
 * SUM/MIN/MAX in the state's type; one row per input row, the batch's
 * selection copied (a final aggregate cannot tell these rows from emitted
 * groups). Columns as agg_finish_rowset's output; kernel-owned. Not
 * available for an aggregate with a fused predicate or computed columns.
 * 0 or -1. */
int qdb_gpu_agg_partial_states(const QdbAggregateKernels* kernels, const TRowSet* batch, TRowSet* out);

/* Decimal storage (DecimalStorageValues, qdb/io/parquet/decimal_storage.cpp:
 * 36-60; DecimalStorageOf, qdb/plan/types/decimal.h:80-99): n unscaled
 * DECIMAL(precision) values of from_width bytes (4, 8, 16 or 32, little-
 * endian two's complement, as Arrow decodes them) at src (device) written to
 * dst (device) in the storage width, 8 bytes for precision <= 18, else 16:
 * the low bytes of the sign-extended value. Asynchronous on the library's
 * stream. 0 or -1. */
int qdb_gpu_decimal_storage(const void* src, int32_t from_width, int64_t n, int32_t precision, void* dst);

/* ORDER BY (TSortProcessor's radix sort, qdb/exec/sort_exec.cpp:393-427;
 * kernels qdb/kernel/sort/radix.oz, radix_nullable.oz): the selected rows
 * of a device batch ordered by the keys, first key most significant, equal
 * keys in input order. Order keys as the reference's qumir_radix_key:
 * signed integers, dates and decimals by value, floats in PostgreSQL order
 * (NaN above +inf, -0 equal to +0), bool false < true; Desc reverses a key;
 * Nulls places its NULL rows (DEFAULT: last ascending, first descending,
 * qdb/plan/ops/sort_key.h:21-26). String keys are not supported. row_index
 * (device, RowCount entries) receives the ordered row indexes, *count how
 * many rows were selected; materialize them with qdb_gpu_gather_rowset
 * (a [start, start + limit) slice of row_index for a LIMIT). 0 or -1. */
#define QDB_SORT_MAX_KEYS 8
enum { QDB_SORT_NULLS_DEFAULT = 0, QDB_SORT_NULLS_FIRST = 1, QDB_SORT_NULLS_LAST = 2 };
typedef struct QdbSortKey {
    int32_t Column;
    int8_t Desc;
    int8_t Nulls;
    int8_t Pad[2];
} QdbSortKey;
int qdb_gpu_sort_rowset(const TRowSet* rs, const QdbColumnSchema* types, const QdbSortKey* keys, int32_t nkeys,
    uint32_t* row_index, int64_t* count);

/* Kernel timing of the last launches, for roofline reporting: the average
 * duration in milliseconds of the named kernel family ("agg_update",
 * "join_probe", ...) measured with HIP events on the stream it runs on since
 * the last reset, and the number of launches. */
void qdb_gpu_timing_enable(int on);
void qdb_gpu_timing_reset(void);
double qdb_gpu_timing_avg_ms(const char* family, int64_t* launches);

/* Sizes of the structs above (TColumn ... QdbFusedResult, in declaration
 * order) so that bindings can check their layouts; returns the count. */
int64_t qdb_gpu_struct_sizes(int64_t* out, int64_t n);

/* Device limits the kernels size themselves by: [compute units, largest LDS
 * allocation of one workgroup in bytes]. Returns the count written, or -1. */
int64_t qdb_gpu_device_limits(int64_t* out, int64_t n);

/* Kernels generated per aggregate (no counterpart in the reference, whose
 * compiler generates its kernels per query; for diagnostics and tests).
 * qdb_gpu_aggregate_kernel_source writes the source of the update kernel
 * generated for `spec` (at most cap-1 bytes and a NUL) and returns its
 * length, or -1; it needs no GPU. qdb_gpu_jit_check compiles a generated
 * source for `arch` without loading it: 0 when it compiles, else -1 with
 * the compiler log in err. qdb_gpu_jit_stats returns how many kernels this
 * process compiled at run time and their total compile time. The
 * environment variable QDB_GPU_JIT=0 runs the precompiled generic kernels
 * instead. */
int64_t qdb_gpu_aggregate_kernel_source(const QdbAggregateSpec* spec, char* out, int64_t cap);
/* The same for a filter: the kernel generated for its predicate (the
 * reference's compiler generates one per filter, qdb/kernel/gen.cpp:
 * 911-1091); its length, or -1. */
int64_t qdb_gpu_filter_kernel_source(const QdbFilterSpec* spec, char* out, int64_t cap);
/* The same for a projection: the kernel generated for its outputs (the
 * reference's compiler generates one per project, qdb/kernel/gen.cpp:
 * 1373-1613); its length, or -1. */
int64_t qdb_gpu_project_kernel_source(const QdbProjectSpec* spec, char* out, int64_t cap);

/* Batched form (no counterpart in the reference, whose processors update
 * one table per call): applies batches[i] to the initialized HashTable blob
 * hts[i], for every i < n, in ONE persistent launch whose workgroups take
 * (table, row range) items from a device work queue. Every table must come
 * from `kernels` (one compiled aggregate). Results are exactly those of n
 * agg_dispatch(Update) calls. Returns 0, or -1 with the error text. */
int qdb_gpu_agg_update_batched(const QdbAggregateKernels* kernels, void* const* hts, const TRowSet* const* batches,
                               int64_t n);
/* The aggregate autotuner (no counterpart in the reference, whose planner
 * picks the table layout by its own rules): an update of a shape bucket
 * (key words, ceil(log2(table slots)), floor(log2(batch rows))) that is not
 * cached yet times every candidate configuration -- one pass over LDS
 * tables of 32, 64 or 128 KiB per workgroup at 4, 2 or 1 workgroups per CU
 * (and 64 KiB with eight copies of the state words, for a few groups), or
 * the bucketed path at 1/2x, 1x or 2x its bucket count with its
 * partition over up to 2048 or 512 workgroups -- as whole
 * queries on scratch tables, caches the fastest, and every later update of
 * that bucket (plain, batched, served) runs in it. Off by default (the
 * sizing rule decides); QDB_GPU_AUTOTUNE=1 turns it on at start-up and
 * QDB_GPU_TUNE_FORCE="path,lds_kib,grid_per_cu,fanout_shift[,part_blocks[,lds_copies]]" forces one
 * configuration for every update. */
typedef struct QdbTuneEntry {
    int32_t KeyWords;
    int32_t CardBucket;
    int32_t RowsBucket;
    int32_t Path;          /* 0: LDS tables + HBM table, 1: bucketed */
    int32_t LdsKiB;        /* path 0 */
    int32_t GridPerCu;     /* path 0 */
    int32_t FanoutShift;   /* path 1: log2 of the bucket count relative to the sizing rule */
    int32_t PartBlocks;    /* path 1: most partition workgroups (0: 2048) */
    int32_t LdsCopies;     /* path 0: copies of the LDS state words (1: none) */
    int32_t Pad;
    double Ms;             /* the chosen configuration's time on the batch it was tuned on */
} QdbTuneEntry;
/* Turns tuning of uncached shapes on (1) or off (0); returns the previous state. */
int qdb_gpu_autotune_enable(int on);
/* Copies up to `max` cache entries into `out` (may be NULL); returns the count. */
int64_t qdb_gpu_autotune_entries(QdbTuneEntry* out, int64_t max);
void qdb_gpu_autotune_clear(void);
/* Writes the cache as text; 0 or -1. */
int qdb_gpu_autotune_save(const char* path);
/* Adds the entries of a file written by qdb_gpu_autotune_save; returns how
 * many were read, or -1. */
int64_t qdb_gpu_autotune_load(const char* path);

/* Batched join build (no counterpart in the reference, whose processors
 * build one table per call): jr_dispatch(Build) of the RowArena / RowTable
 * blobs arenas[i] / tables[i], for every i < n, all of `kernels` (one
 * compiled join) and appended to. The dense tables' kernels and readbacks of
 * every join are queued before ONE synchronization; a join whose keys the
 * dense table does not take is built with the hash table as by its own
 * Build. Returns 0, or -1 with the error text. */
int qdb_gpu_join_build_batched(const QdbRowJoinKernels* kernels, void* const* arenas, void* const* tables, int64_t n);
/* Detached batched probe (no counterpart in the reference): batches[i]
 * (i < n) probed against ONE built join `arena` (one-word keys), each into
 * its own pair buffer pairs[i] (pairs NULL, or pairs[i] NULL: the pairs are
 * counted, not written), with the pair count in counts[i] and the fused
 * results of that batch alone in fused[i * NumFused + f] (fused may be
 * NULL). The join's own pair buffer, counters and fused totals are not
 * touched, so threads that each issue on their own stream
 * (qdb_gpu_set_thread_stream) may probe one resident build table at the
 * same time. Each call replays a hipGraph captured once per stream and
 * kernel mode (queue reset, counter set-up, the persistent probe over
 * (batch, row range) items, collection of the counts and fused words, one
 * copy to pinned host memory) after one descriptor upload, up to 256
 * batches per replay. Returns 0, or -1 with the error text. */
int qdb_gpu_join_probe_detached(const QdbRowJoinKernels* kernels, void* arena, const TRowSet* const* batches,
                                QdbPairBuffer* const* pairs, int64_t n, int64_t* counts, QdbFusedResult* fused);
/* Diagnostics of a join's table (no counterpart in the reference): [built,
 * dense direct table, dense key range, bits per packed ordinal (0: not
 * packed), rows kept, rows appended]; returns how many values were written,
 * or -1. */
int64_t qdb_gpu_join_table_info(void* arena, int64_t* out, int64_t n);
/* Makes the calling thread issue its work on library stream `index`
 * (0 = the default library stream, up to 3 more, all at the default
 * priority). Device blocks freed on one stream are reused by another only
 * after the work queued before the free has completed. One thread per
 * stream index at a time. 0 or -1. */
int qdb_gpu_set_thread_stream(int index);
/* Order-independent checksum of n pairs at `pairs` (device, 2 i64 each):
 * out[0] = sum and out[1] = xor of a 64-bit mix of every pair (no
 * counterpart in the reference: it compares pair sets without moving them
 * to the host). 0 or -1. */
int qdb_gpu_pairs_checksum(const int64_t* pairs, int64_t n, uint64_t* out);
/* Batched join probe (no counterpart in the reference): jr_dispatch(Probe)
 * of batches[i] against the built join arenas[i] into the PairBuffer blob
 * pairs[i], for every i < n, in ONE persistent launch whose workgroups take
 * (join, row range) items from a device work queue (joins with one-word
 * keys; two-word keys probe one join at a time). Pair buffers, their Count
 * and the fused results (qdb_gpu_join_fused_results) are as after n
 * separate probes; pairs[i] is materialized with batches[i]. Returns 0, or
 * -1 with the error text. */
int qdb_gpu_join_probe_batched(const QdbRowJoinKernels* kernels, void* const* arenas, const TRowSet* const* batches,
                               void* const* pairs, int64_t n);
/* Batched finish (no counterpart in the reference): agg_finish_rowset of
 * hts[i] into outs[i] for every i < n, every table's finalize launched
 * before ONE synchronization; rows[i] gets outs[i]'s row count and each
 * table is destroyed, as by agg_finish_rowset. Returns 0, or -1 with the
 * error text (outputs already made are released). */
int qdb_gpu_agg_finish_batched(const QdbAggregateKernels* kernels, void* const* hts, TRowSet* outs, int64_t* rows,
                               int64_t n);
int qdb_gpu_jit_check(const char* src, const char* arch, char* err, int64_t cap);
int64_t qdb_gpu_jit_stats(double* total_ms);

/* ------------------------------------------------------------------------ */
/* Host drivers over the entry points above, for callers without the       */
/* reference's processors: they drive the same op-code protocol and check   */
/* the same invariants as TAggregateProcessor (qdb/exec/aggregate_exec.cpp: */
/* 67-136: Init once, Update < 0 fails, finalized rows == Size read before  */
/* the finish, the output adopted) and TRowJoinProcessor                    */
/* (qdb/exec/row_join_exec.cpp:167-190: pairs materialized before the probe */
/* batch is released, pair count reset per drain). Batches passed in stay   */
/* owned by the caller; outputs are adopted (Destroy set). Errors: -1 / NULL,*/
/* text in qdb_gpu_exec_last_error.                                          */
/* ------------------------------------------------------------------------ */

const char* qdb_gpu_exec_last_error(void);
/* processors[i] (all of one compiled aggregate) take batches[i], in one
 * batched launch (qdb_gpu_agg_update_batched). 0 or -1. */
int qdb_gpu_agg_processor_add_batched(void* const* procs, const TRowSet* const* batches, int64_t n);
/* processors[i] (all of one compiled aggregate) finish into outs[i], with
 * one synchronization (qdb_gpu_agg_finish_batched). 0 or -1. */
int qdb_gpu_agg_processor_finish_batched(void* const* procs, TRowSet* outs, int64_t n);
/* row_layout != 0 starts with InitRows(initial_groups), else Init. */
void* qdb_gpu_agg_processor_new(const QdbAggregateKernels* kernels, int64_t initial_capacity, int row_layout,
    int64_t initial_groups);
int qdb_gpu_agg_processor_add(void* processor, const TRowSet* batch);
/* Finalizes into `out` (release it with qdb_gpu_rowset_release); returns rows. */
int64_t qdb_gpu_agg_processor_finish(void* processor, TRowSet* out);
int64_t qdb_gpu_agg_processor_groups(void* processor);
void qdb_gpu_agg_processor_free(void* processor);

void* qdb_gpu_join_processor_new(const QdbRowJoinKernels* kernels, int build_is_left, int64_t output_batch_rows);
int qdb_gpu_join_processor_append(void* processor, const TRowSet* batch);
int qdb_gpu_join_processor_build(void* processor);
/* Probes and materializes the batch's matches into queued output batches;
 * returns the pair count. */
int64_t qdb_gpu_join_processor_probe(void* processor, const TRowSet* batch);
/* Probes only; returns the pair count, pairs readable until the next probe. */
int64_t qdb_gpu_join_processor_probe_only(void* processor, const TRowSet* batch);
const QdbPairBuffer* qdb_gpu_join_processor_pairs(void* processor);
/* Builds of processors[i] (all of one compiled join) with one
 * synchronization (qdb_gpu_join_build_batched). 0 or -1. */
int qdb_gpu_join_processor_build_batched(void* const* processors, int64_t n);
/* probe_only of batches[i] by processors[i] (all of one compiled join) in
 * one persistent launch (qdb_gpu_join_probe_batched). 0 or -1. */
int qdb_gpu_join_processor_probe_batched(void* const* processors, const TRowSet* const* batches, int64_t n);
/* The processor's RowArena blob (for qdb_gpu_join_fused_results). */
void* qdb_gpu_join_processor_arena(void* processor);
/* Next queued output batch: 1 when one was returned, 0 when none is left. */
int qdb_gpu_join_processor_next(void* processor, TRowSet* out);
void qdb_gpu_join_processor_free(void* processor);
/* The partial aggregate (TPartialAggregateProcessor, qdb/exec/
 * aggregate_exec.cpp:158-233; TPartialAggregatePolicy, qdb/exec/
 * plan_data.h:234-250, defaults 131072 / 8 / 2^20): add consumes a batch;
 * the table emits its groups into a pending output when it reaches
 * max_groups and restarts; with pass_through, once sample_rows rows were
 * hashed at fewer than min_reduction rows per produced group, the table is
 * emitted and every later batch comes out as its partial states
 * (qdb_gpu_agg_partial_states). next returns 1 with a pending output
 * (adopted), 0 when none; finish returns the table's last groups' row count,
 * 0 when rows were passed through (no table remains), -1 on error. */
void* qdb_gpu_partial_agg_new(const QdbAggregateKernels* kernels, int64_t initial_capacity, uint64_t sample_rows,
    uint64_t min_reduction, int64_t max_groups, int pass_through);
int qdb_gpu_partial_agg_add(void* processor, const TRowSet* batch);
int qdb_gpu_partial_agg_next(void* processor, TRowSet* out);
int64_t qdb_gpu_partial_agg_finish(void* processor, TRowSet* out);
int qdb_gpu_partial_agg_passing_through(void* processor);
void qdb_gpu_partial_agg_free(void* processor);
/* Calls rs->Destroy when set, else qdb_gpu_destroy_kernel_rowset. */
void qdb_gpu_rowset_release(TRowSet* rs);

#ifdef __cplusplus
}
#endif

#endif /* QDB_GPU_H */
