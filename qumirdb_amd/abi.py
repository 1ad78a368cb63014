"""ctypes mirror of include/qdb_gpu.h (the C-ABI of the HIP library).

Struct layouts follow the header field by field; tests compare every size
with the library's own sizeof (qdb_gpu_struct_sizes).
"""
from __future__ import annotations

import ctypes as C

MAX_INSNS = 32
MAX_REGS = 8
MAX_COLS = 64
PROJECT_MAX_OUTPUTS = 8
AGG_MAX_KEYS = 4
AGG_MAX_AGGS = 16
AGG_MAX_EMPTY_INPUT_GROUPS = 4
JOIN_MAX_KEYS = 2
JOIN_MAX_FUSED = 4

# Value kinds (qdb/io/value_type.h:7-18)
SIGNED, UNSIGNED, FLOAT, BOOL, STRING, DATE, DECIMAL = range(7)

# Expression ops
(EXPR_COLUMN, EXPR_CONST_INT, EXPR_CONST_F64, EXPR_ADD, EXPR_SUB, EXPR_MUL, EXPR_NEG, EXPR_SCALE_UP,
 EXPR_CMP, EXPR_AND, EXPR_OR, EXPR_NOT, EXPR_IS_NULL, EXPR_TO_F64, EXPR_DIV) = range(1, 16)
CMP_EQ, CMP_NE, CMP_LT, CMP_LE, CMP_GT, CMP_GE = range(6)
REG_INT, REG_F64, REG_BOOL = range(3)

# Aggregate functions and op-codes
AGG_COUNT_STAR, AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX = range(5)
AGG_OP_INIT, AGG_OP_UPDATE, AGG_OP_INIT_ROWS, AGG_OP_SWITCH_TO_ROWS_AT, AGG_OP_EXPECT_ROW_GROUPS = 0, 1, 3, 4, 5
JR_OP_INIT, JR_OP_APPEND, JR_OP_BUILD, JR_OP_PROBE = 0, 1, 2, 3
JOIN_INNER, JOIN_LEFT_OUTER, JOIN_RIGHT_OUTER = range(3)
JOIN_TABLE_AUTO, JOIN_TABLE_HASH = range(2)
SLOT_FILTER, SLOT_PROJECT, SLOT_AGGREGATE, SLOT_JOIN, SLOT_HASH = range(5)

HASH_TABLE_SIZE = 304
PAIR_BUFFER_SIZE = 24
ROW_ARENA_SIZE = 56
ROW_TABLE_SIZE = 48
ROW_SCRATCH_SIZE = 64


class TColumn(C.Structure):
    _fields_ = [
        ("Data", C.c_void_p),
        ("DataBitOffset", C.c_int32),
        ("Mask", C.c_void_p),
        ("MaskBitOffset", C.c_int32),
        ("Offsets", C.c_void_p),
        ("OffsetWidth", C.c_uint8),
    ]


class TRowSet(C.Structure):
    pass


DESTROY_FN = C.CFUNCTYPE(None, C.POINTER(TRowSet))

TRowSet._fields_ = [
    ("Columns", C.POINTER(TColumn)),
    ("ColumnCount", C.c_int64),
    ("RowCount", C.c_int64),
    ("Selection", C.c_void_p),
    ("Hash", C.c_void_p),
    ("Destroy", C.c_void_p),
    ("Private", C.c_void_p),
    ("RefCount", C.c_int64),
]


class QdbValueType(C.Structure):
    _fields_ = [("Kind", C.c_uint8), ("Width", C.c_int32), ("Scale", C.c_int32), ("Nullable", C.c_bool),
                ("Precision", C.c_uint8)]


class QdbExprInsn(C.Structure):
    _fields_ = [
        ("Op", C.c_uint8), ("Dst", C.c_uint8), ("A", C.c_uint8), ("B", C.c_uint8), ("Class", C.c_uint8),
        ("Pad", C.c_uint8 * 3), ("Column", C.c_int32), ("Pad2", C.c_int32), ("ImmLo", C.c_int64), ("ImmHi", C.c_int64),
    ]


class QdbExprProgram(C.Structure):
    _fields_ = [("NumInsns", C.c_int32), ("Result", C.c_int32), ("Insns", QdbExprInsn * MAX_INSNS)]


class QdbColumnSchema(C.Structure):
    _fields_ = [("NumColumns", C.c_int32), ("Types", QdbValueType * MAX_COLS)]


SORT_MAX_KEYS = 8
SORT_NULLS_DEFAULT, SORT_NULLS_FIRST, SORT_NULLS_LAST = 0, 1, 2


class QdbSortKey(C.Structure):
    _fields_ = [("Column", C.c_int32), ("Desc", C.c_int8), ("Nulls", C.c_int8), ("Pad", C.c_int8 * 2)]


class QdbFilterSpec(C.Structure):
    _fields_ = [("Input", QdbColumnSchema), ("Predicate", QdbExprProgram)]


class QdbProjectSpec(C.Structure):
    _fields_ = [
        ("Input", QdbColumnSchema),
        ("NumOutputs", C.c_int32),
        ("OutputTypes", QdbValueType * PROJECT_MAX_OUTPUTS),
        ("Outputs", QdbExprProgram * PROJECT_MAX_OUTPUTS),
    ]


class QdbAggregateFn(C.Structure):
    _fields_ = [("Func", C.c_int32), ("Column", C.c_int32)]


class QdbAggregateSpec(C.Structure):
    _fields_ = [
        ("Input", QdbColumnSchema),
        ("NumKeys", C.c_int32),
        ("KeyColumns", C.c_int32 * AGG_MAX_KEYS),
        ("NumAggs", C.c_int32),
        ("Aggs", QdbAggregateFn * AGG_MAX_AGGS),
        ("NumEmptyInputGroups", C.c_int32),
        ("EmptyInputGroups", C.c_int64 * AGG_MAX_EMPTY_INPUT_GROUPS),
        ("OutputOnHost", C.c_int32),
        ("HasPredicate", C.c_int32),
        ("Predicate", QdbExprProgram),
        ("NumComputed", C.c_int32),
        ("ComputedTypes", QdbValueType * PROJECT_MAX_OUTPUTS),
        ("Computed", QdbExprProgram * PROJECT_MAX_OUTPUTS),
    ]


AGG_DISPATCH_FN = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.POINTER(TRowSet), C.c_int64, C.c_int64)
AGG_FINISH_FN = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.POINTER(TRowSet))
FILTER_FN = C.CFUNCTYPE(None, C.POINTER(TRowSet), C.c_void_p, C.c_void_p, C.c_void_p)
PROJECT_FN = C.CFUNCTYPE(None, C.POINTER(TRowSet), C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p)
JOIN_DISPATCH_FN = C.CFUNCTYPE(C.c_bool, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(TRowSet),
                               C.POINTER(C.c_int64), C.c_int64)
JOIN_MATERIALIZE_FN = C.CFUNCTYPE(C.c_int64, C.c_void_p, C.POINTER(TRowSet), C.c_int64, C.c_int64, C.POINTER(TRowSet))
GROUP_HASH_FN = C.CFUNCTYPE(C.c_bool, C.POINTER(TRowSet), C.c_void_p, C.POINTER(C.c_int64))
HEAP_ALLOC_FN = C.CFUNCTYPE(C.c_void_p, C.c_int64)
HEAP_FREE_FN = C.CFUNCTYPE(None, C.c_void_p)
HEAP_DETACH_FN = C.CFUNCTYPE(None, C.c_void_p)
ROW_ERROR_FN = C.CFUNCTYPE(C.c_int64, C.c_int64)


class QdbAggregateKernels(C.Structure):
    _fields_ = [
        ("Dispatch", C.c_void_p),
        ("FinishRowSet", C.c_void_p),
        ("NumAggs", C.c_int64),
        ("StateWords", C.c_int64),
        ("RowBytes", C.c_int64),
        ("Slot", C.c_int32),
        ("NumOutputColumns", C.c_int32),
        ("OutputTypes", QdbValueType * (AGG_MAX_KEYS + AGG_MAX_AGGS)),
    ]


class QdbRowJoinSpec(C.Structure):
    _fields_ = [
        ("Build", QdbColumnSchema),
        ("Probe", QdbColumnSchema),
        ("NumKeys", C.c_int32),
        ("BuildKeyColumns", C.c_int32 * JOIN_MAX_KEYS),
        ("ProbeKeyColumns", C.c_int32 * JOIN_MAX_KEYS),
        ("JoinType", C.c_int32),
        ("ProbeIsLeft", C.c_int32),
        ("NumFused", C.c_int32),
        ("Fused", QdbAggregateFn * JOIN_MAX_FUSED),
        ("OutputOnHost", C.c_int32),
        ("TableMode", C.c_int32),
    ]


class QdbRowJoinKernels(C.Structure):
    _fields_ = [
        ("Dispatch", C.c_void_p),
        ("Materialize", C.c_void_p),
        ("Slot", C.c_int32),
        ("NumOutputColumns", C.c_int32),
        ("OutputTypes", QdbValueType * 128),
    ]


class QdbPairBuffer(C.Structure):
    _fields_ = [("Count", C.c_int64), ("Capacity", C.c_int64), ("Data", C.c_void_p)]


class QdbFusedResult(C.Structure):
    _fields_ = [("Lo", C.c_int64), ("Hi", C.c_int64), ("Count", C.c_int64), ("Valid", C.c_int64)]


class QdbTuneEntry(C.Structure):
    _fields_ = [("KeyWords", C.c_int32), ("CardBucket", C.c_int32), ("RowsBucket", C.c_int32), ("Path", C.c_int32),
                ("LdsKiB", C.c_int32), ("GridPerCu", C.c_int32), ("FanoutShift", C.c_int32), ("PartBlocks", C.c_int32), ("LdsCopies", C.c_int32), ("Pad", C.c_int32),
                ("Ms", C.c_double)]


# Every struct the header declares, in the order qdb_gpu_struct_sizes reports.
STRUCTS = [TColumn, TRowSet, QdbValueType, QdbExprInsn, QdbExprProgram, QdbColumnSchema, QdbFilterSpec,
           QdbProjectSpec, QdbAggregateFn, QdbAggregateSpec, QdbAggregateKernels, QdbRowJoinSpec,
           QdbRowJoinKernels, QdbPairBuffer, QdbFusedResult, QdbTuneEntry]

# Every extern "C" symbol include/qdb_gpu.h declares.
EXPORTS = [
    "qdb_gpu_init", "qdb_gpu_last_error", "qdb_gpu_version",
    "qdb_gpu_compile_filter", "qdb_gpu_compile_project", "qdb_gpu_compile_aggregate", "qdb_gpu_compile_row_join",
    "qdb_gpu_compile_group_hash", "qdb_gpu_release_slot", "qdb_gpu_join_fused_results", "qdb_gpu_join_destroy",
    "qdb_gpu_aggregate_destroy", "qdb_gpu_alloc", "qdb_gpu_free", "qdb_gpu_memcpy_h2d", "qdb_gpu_memcpy_d2h",
    "qdb_gpu_memset", "qdb_gpu_synchronize", "qdb_gpu_stage_rowset", "qdb_gpu_download_rowset",
    "qdb_gpu_gen_lineitem", "qdb_gpu_gen_join_build", "qdb_gpu_gen_join_probe", "qdb_gpu_gen_zipf_groups",
    "qdb_gpu_timing_enable", "qdb_gpu_timing_reset", "qdb_gpu_timing_avg_ms", "qdb_gpu_struct_sizes",
    "qdb_gpu_device_limits", "qdb_gpu_aggregate_kernel_source", "qdb_gpu_jit_check", "qdb_gpu_jit_stats",
    "qdb_gpu_agg_update_batched", "qdb_gpu_agg_processor_add_batched", "qdb_gpu_agg_finish_batched",
    "qdb_gpu_filter_kernel_source", "qdb_gpu_project_kernel_source",
    "qdb_gpu_agg_processor_finish_batched", "qdb_gpu_trim_cache", "qdb_gpu_pool_stats",
    "qdb_gpu_exec_last_error", "qdb_gpu_agg_processor_new", "qdb_gpu_agg_processor_add",
    "qdb_gpu_agg_processor_finish", "qdb_gpu_agg_processor_groups", "qdb_gpu_agg_processor_free",
    "qdb_gpu_join_processor_new", "qdb_gpu_join_processor_append", "qdb_gpu_join_processor_build",
    "qdb_gpu_join_processor_probe", "qdb_gpu_join_processor_probe_only", "qdb_gpu_join_processor_pairs",
    "qdb_gpu_join_processor_arena", "qdb_gpu_join_processor_next", "qdb_gpu_join_processor_free",
    "qdb_gpu_rowset_release", "qdb_gpu_set_host_heap", "qdb_gpu_destroy_kernel_rowset", "qdb_gpu_adopt_kernel_rowset",
    "qdb_gpu_set_row_error_handler", "qdb_gpu_take_row_error", "qdb_gpu_join_probe_fused_noalloc",
    "qdb_gpu_decimal_storage", "qdb_gpu_sort_rowset", "qdb_gpu_shuffle_partition", "qdb_gpu_gather_rowset", "qdb_gpu_agg_partial_states", "qdb_gpu_partial_agg_new",
    "qdb_gpu_partial_agg_add", "qdb_gpu_partial_agg_next", "qdb_gpu_partial_agg_finish", "qdb_gpu_partial_agg_passing_through",
    "qdb_gpu_partial_agg_free",
    "qdb_gpu_rf_new", "qdb_gpu_rf_build", "qdb_gpu_rf_publish", "qdb_gpu_rf_probe", "qdb_gpu_rf_stats", "qdb_gpu_rf_free",
    "qdb_gpu_join_build_batched", "qdb_gpu_join_probe_batched", "qdb_gpu_join_processor_build_batched",
    "qdb_gpu_join_processor_probe_batched", "qdb_gpu_pairs_checksum",
    "qdb_gpu_autotune_enable", "qdb_gpu_autotune_entries", "qdb_gpu_autotune_clear", "qdb_gpu_autotune_save",
    "qdb_gpu_autotune_load", "qdb_gpu_join_probe_detached", "qdb_gpu_set_thread_stream", "qdb_gpu_join_table_info",
]


def declare(lib: C.CDLL) -> C.CDLL:
    """Sets argument and result types of the library's C entry points."""
    P = C.c_void_p
    i64 = C.c_int64
    sigs = {
        "qdb_gpu_init": (C.c_int, [C.c_int]),
        "qdb_gpu_last_error": (C.c_char_p, []),
        "qdb_gpu_version": (C.c_char_p, []),
        "qdb_gpu_compile_filter": (C.c_int, [C.POINTER(QdbFilterSpec), C.POINTER(P)]),
        "qdb_gpu_compile_project": (C.c_int, [C.POINTER(QdbProjectSpec), C.POINTER(P)]),
        "qdb_gpu_compile_aggregate": (C.c_int, [C.POINTER(QdbAggregateSpec), C.POINTER(QdbAggregateKernels)]),
        "qdb_gpu_compile_row_join": (C.c_int, [C.POINTER(QdbRowJoinSpec), C.POINTER(QdbRowJoinKernels)]),
        "qdb_gpu_compile_group_hash": (C.c_int, [C.POINTER(QdbColumnSchema), C.c_int32, C.POINTER(P)]),
        "qdb_gpu_release_slot": (None, [C.c_int32, C.c_int32]),
        "qdb_gpu_join_fused_results": (i64, [P, C.POINTER(QdbFusedResult), i64]),
        "qdb_gpu_join_destroy": (None, [P, P, P, P]),
        "qdb_gpu_aggregate_destroy": (None, [P]),
        "qdb_gpu_alloc": (P, [C.c_size_t]),
        "qdb_gpu_free": (None, [P]),
        "qdb_gpu_memcpy_h2d": (C.c_int, [P, P, C.c_size_t]),
        "qdb_gpu_memcpy_d2h": (C.c_int, [P, P, C.c_size_t]),
        "qdb_gpu_memset": (C.c_int, [P, C.c_int, C.c_size_t]),
        "qdb_gpu_synchronize": (C.c_int, []),
        "qdb_gpu_stage_rowset": (C.c_int, [C.POINTER(TRowSet), C.POINTER(QdbColumnSchema), C.POINTER(TRowSet)]),
        "qdb_gpu_download_rowset": (C.c_int, [C.POINTER(TRowSet), C.POINTER(QdbValueType), C.POINTER(TRowSet)]),
        "qdb_gpu_gen_lineitem": (C.c_int, [C.c_uint64, i64, i64, P, P, P, P, P, P, P]),
        "qdb_gpu_gen_join_build": (C.c_int, [C.c_uint64, i64, i64, P, P]),
        "qdb_gpu_gen_join_probe": (C.c_int, [C.c_uint64, i64, i64, i64, P, P]),
        "qdb_gpu_gen_zipf_groups": (C.c_int, [C.c_uint64, i64, i64, P, i64, P, P, P]),
        "qdb_gpu_timing_enable": (None, [C.c_int]),
        "qdb_gpu_timing_reset": (None, []),
        "qdb_gpu_timing_avg_ms": (C.c_double, [C.c_char_p, C.POINTER(i64)]),
        "qdb_gpu_device_limits": (i64, [C.POINTER(i64), i64]),
        "qdb_gpu_aggregate_kernel_source": (i64, [C.POINTER(QdbAggregateSpec), C.c_char_p, i64]),
        "qdb_gpu_jit_check": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, i64]),
        "qdb_gpu_jit_stats": (i64, [C.POINTER(C.c_double)]),
        "qdb_gpu_agg_update_batched": (C.c_int, [C.POINTER(QdbAggregateKernels), C.POINTER(P), C.POINTER(C.POINTER(TRowSet)), i64]),
        "qdb_gpu_agg_processor_add_batched": (C.c_int, [C.POINTER(P), C.POINTER(C.POINTER(TRowSet)), i64]),
        "qdb_gpu_agg_finish_batched": (C.c_int, [C.POINTER(QdbAggregateKernels), C.POINTER(P), C.POINTER(TRowSet),
                                                 C.POINTER(i64), i64]),
        "qdb_gpu_agg_processor_finish_batched": (C.c_int, [C.POINTER(P), C.POINTER(TRowSet), i64]),
        "qdb_gpu_filter_kernel_source": (i64, [C.POINTER(QdbFilterSpec), C.c_char_p, i64]),
        "qdb_gpu_project_kernel_source": (i64, [C.POINTER(QdbProjectSpec), C.c_char_p, i64]),
        "qdb_gpu_trim_cache": (None, []),
        "qdb_gpu_pool_stats": (i64, [P, i64]),
        "qdb_gpu_struct_sizes": (i64, [C.POINTER(i64), i64]),
        "qdb_gpu_exec_last_error": (C.c_char_p, []),
        "qdb_gpu_agg_processor_new": (P, [C.POINTER(QdbAggregateKernels), i64, C.c_int, i64]),
        "qdb_gpu_agg_processor_add": (C.c_int, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_agg_processor_finish": (i64, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_agg_processor_groups": (i64, [P]),
        "qdb_gpu_agg_processor_free": (None, [P]),
        "qdb_gpu_join_processor_new": (P, [C.POINTER(QdbRowJoinKernels), C.c_int, i64]),
        "qdb_gpu_join_processor_append": (C.c_int, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_join_processor_build": (C.c_int, [P]),
        "qdb_gpu_join_processor_probe": (i64, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_join_processor_probe_only": (i64, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_join_processor_pairs": (C.POINTER(QdbPairBuffer), [P]),
        "qdb_gpu_join_processor_arena": (P, [P]),
        "qdb_gpu_join_processor_next": (C.c_int, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_join_processor_free": (None, [P]),
        "qdb_gpu_rowset_release": (None, [C.POINTER(TRowSet)]),
        "qdb_gpu_set_host_heap": (None, [HEAP_ALLOC_FN, HEAP_FREE_FN, HEAP_DETACH_FN]),
        "qdb_gpu_destroy_kernel_rowset": (None, [C.POINTER(TRowSet)]),
        "qdb_gpu_adopt_kernel_rowset": (None, [C.POINTER(TRowSet)]),
        "qdb_gpu_set_row_error_handler": (None, [ROW_ERROR_FN]),
        "qdb_gpu_take_row_error": (i64, []),
        "qdb_gpu_join_probe_fused_noalloc": (C.c_int, [P, C.POINTER(TRowSet), C.POINTER(QdbFusedResult), i64,
                                                       C.POINTER(i64)]),
        "qdb_gpu_decimal_storage": (C.c_int, [P, C.c_int32, i64, C.c_int32, P]),
        "qdb_gpu_sort_rowset": (C.c_int, [C.POINTER(TRowSet), C.POINTER(QdbColumnSchema), C.POINTER(QdbSortKey), C.c_int32, P,
                                         C.POINTER(i64)]),
        "qdb_gpu_shuffle_partition": (C.c_int, [C.POINTER(TRowSet), P, C.c_int32, P, C.c_int32, P, C.POINTER(i64)]),
        "qdb_gpu_gather_rowset": (C.c_int, [C.POINTER(TRowSet), C.POINTER(QdbColumnSchema), P, i64, C.POINTER(TRowSet)]),
        "qdb_gpu_agg_partial_states": (C.c_int, [C.POINTER(QdbAggregateKernels), C.POINTER(TRowSet), C.POINTER(TRowSet)]),
        "qdb_gpu_partial_agg_new": (P, [C.POINTER(QdbAggregateKernels), i64, C.c_uint64, C.c_uint64, i64, C.c_int]),
        "qdb_gpu_partial_agg_add": (C.c_int, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_partial_agg_next": (C.c_int, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_partial_agg_finish": (i64, [P, C.POINTER(TRowSet)]),
        "qdb_gpu_partial_agg_passing_through": (C.c_int, [P]),
        "qdb_gpu_partial_agg_free": (None, [P]),
        "qdb_gpu_rf_new": (P, [i64, i64]),
        "qdb_gpu_rf_build": (C.c_int, [P, P, P, i64]),
        "qdb_gpu_rf_publish": (C.c_int, [P]),
        "qdb_gpu_rf_probe": (i64, [P, P, P, P, i64]),
        "qdb_gpu_rf_stats": (i64, [P, C.POINTER(i64), i64]),
        "qdb_gpu_rf_free": (None, [P]),
        "qdb_gpu_join_build_batched": (C.c_int, [C.POINTER(QdbRowJoinKernels), C.POINTER(P), C.POINTER(P), i64]),
        "qdb_gpu_join_probe_batched": (C.c_int, [C.POINTER(QdbRowJoinKernels), C.POINTER(P), C.POINTER(C.POINTER(TRowSet)),
                                                 C.POINTER(P), i64]),
        "qdb_gpu_join_processor_build_batched": (C.c_int, [C.POINTER(P), i64]),
        "qdb_gpu_join_processor_probe_batched": (C.c_int, [C.POINTER(P), C.POINTER(C.POINTER(TRowSet)), i64]),
        "qdb_gpu_pairs_checksum": (C.c_int, [P, i64, C.POINTER(C.c_uint64)]),
        "qdb_gpu_autotune_enable": (C.c_int, [C.c_int]),
        "qdb_gpu_autotune_entries": (i64, [C.POINTER(QdbTuneEntry), i64]),
        "qdb_gpu_autotune_clear": (None, []),
        "qdb_gpu_autotune_save": (C.c_int, [C.c_char_p]),
        "qdb_gpu_autotune_load": (i64, [C.c_char_p]),
        "qdb_gpu_join_probe_detached": (C.c_int, [C.POINTER(QdbRowJoinKernels), P, C.POINTER(C.POINTER(TRowSet)),
                                                  C.POINTER(C.POINTER(QdbPairBuffer)), i64, C.POINTER(i64),
                                                  C.POINTER(QdbFusedResult)]),
        "qdb_gpu_set_thread_stream": (C.c_int, [C.c_int]),
        "qdb_gpu_join_table_info": (i64, [P, C.POINTER(i64), i64]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib
