"""ctypes wrapper of the CPU oracle (oracle.c). TEST INFRASTRUCTURE ONLY: used
by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.

The spec structs are the C-ABI's (qumirdb_amd/abi.py), loaded from their file
so that this module does not load the HIP library.
"""
from __future__ import annotations

import ctypes as C
import importlib.util
import os
import subprocess
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "liboracle.so")


def _load_abi():
    """The ABI module, registered under its package name so that the HIP
    bindings (imported before or after) share the very same struct classes;
    loading it by file does not load the HIP library."""
    import sys
    name = "qumirdb_amd.abi"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "qumirdb_amd", "abi.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


abi = _load_abi()


def build() -> str:
    """Compiles oracle.c (gcc, by oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def _lib() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    P, i64 = C.c_void_p, C.c_int64
    sigs = {
        "oracle_decimal_to_f64": (C.c_double, [i64, i64, i64]),
        "oracle_decimal64_to_f64": (C.c_double, [i64, i64]),
        "oracle_exact_sum": (C.c_double, [P, i64]),
        "oracle_kahan_sum": (C.c_double, [P, i64]),
        "oracle_filter": (None, [C.POINTER(abi.QdbFilterSpec), C.POINTER(abi.TRowSet), P, P]),
        "oracle_rf_mix": (C.c_uint64, [C.c_uint64]),
        "oracle_rf_new": (P, [i64, i64]),
        "oracle_rf_build": (None, [P, P, P, i64]),
        "oracle_rf_publish": (C.c_int, [P]),
        "oracle_rf_probe": (i64, [P, P, P, P, i64]),
        "oracle_rf_free": (None, [P]),
        "oracle_project": (C.c_int, [C.POINTER(abi.QdbProjectSpec), C.POINTER(abi.TRowSet), C.POINTER(P), P]),
        "oracle_key_hash": (C.c_uint64, [C.POINTER(abi.TColumn), C.POINTER(abi.QdbValueType), C.POINTER(C.c_int32), C.c_int, i64]),
        "oracle_mix_u64": (C.c_uint64, [C.c_uint64]),
        "oracle_agg_new": (P, [C.POINTER(abi.QdbAggregateSpec), i64, C.c_int]),
        "oracle_agg_update": (None, [P, C.POINTER(abi.TRowSet)]),
        "oracle_agg_size": (i64, [P]),
        "oracle_agg_finish": (i64, [P, C.POINTER(abi.TRowSet)]),
        "oracle_agg_free": (None, [P]),
        "oracle_agg_output_types": (None, [C.POINTER(abi.QdbAggregateSpec), C.POINTER(abi.QdbValueType)]),
        "oracle_rowset_release": (None, [C.POINTER(abi.TRowSet)]),
        "oracle_join_new": (P, [C.POINTER(abi.QdbRowJoinSpec)]),
        "oracle_join_free": (None, [P]),
        "oracle_join_append": (None, [P, C.POINTER(abi.TRowSet)]),
        "oracle_join_build": (None, [P]),
        "oracle_join_probe": (i64, [P, C.POINTER(abi.TRowSet)]),
        "oracle_join_probe_parallel": (i64, [P, C.POINTER(abi.TRowSet), C.c_int]),
        "oracle_join_pairs": (C.POINTER(C.c_int64), [P]),
        "oracle_join_keys": (i64, [P]),
        "oracle_join_rows": (i64, [P]),
        "oracle_gen_lineitem": (None, [C.c_uint64, i64, i64, P, P, P, P, P, P, P]),
        "oracle_gen_join_build": (None, [C.c_uint64, i64, i64, i64, P, P]),
        "oracle_gen_join_probe": (None, [C.c_uint64, i64, i64, i64, P, P]),
        "oracle_zipf_cdf": (None, [P, i64]),
        "oracle_gen_groups": (None, [C.c_uint64, i64, i64, P, i64, P, P, P]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


LIB = _lib()


def _np_dtype(t):
    kind, width = t.kind, t.width
    if kind == abi.FLOAT:
        return np.float64 if width == 8 else np.float32
    if width == 16:
        return np.dtype([("lo", "<u8"), ("hi", "<i8")])
    signed = kind != abi.UNSIGNED
    return {1: np.int8 if signed else np.uint8, 2: np.int16 if signed else np.uint16,
            4: np.int32 if signed else np.uint32, 8: np.int64 if signed else np.uint64}[width]


def _i128_array(values) -> np.ndarray:
    dt = np.dtype([("lo", "<u8"), ("hi", "<i8")])
    out = np.empty(len(values), dtype=dt)
    for i, v in enumerate(values):
        u = int(v or 0) & ((1 << 128) - 1)
        out[i] = (u & ((1 << 64) - 1), (u >> 64) - (1 << 64) if (u >> 127) else (u >> 64))
    return out


class HostRowSet:
    """A TRowSet over host numpy buffers (kept alive by this object)."""

    def __init__(self, types, columns, valid=None, selection=None, hashes=None):
        self.types = list(types)
        self.keep = []
        self.columns = list(columns)
        self.valid = list(valid) if valid is not None else [None] * len(self.types)
        self.selection = selection
        n = None
        cols = (abi.TColumn * max(1, len(self.types)))()
        for i, (t, data) in enumerate(zip(self.types, columns)):
            if t.kind == abi.BOOL:
                arr = np.packbits(np.asarray(data).astype(np.uint8), bitorder="little")
                n = len(data) if n is None else n
            elif t.width == 16 and not (isinstance(data, np.ndarray) and data.dtype.names):
                arr = _i128_array(data)
                n = len(data) if n is None else n
            else:
                arr = np.ascontiguousarray(np.asarray(data, dtype=_np_dtype(t)))
                n = len(arr) if n is None else n
            arr = np.concatenate([arr.view(np.uint8), np.zeros(16, np.uint8)])  # room for bit reads
            self.keep.append(arr)
            cols[i].Data = arr.ctypes.data
            v = None if valid is None else valid[i]
            if v is not None:
                m = np.concatenate([np.packbits(np.asarray(v, dtype=np.uint8), bitorder="little"), np.zeros(8, np.uint8)])
                self.keep.append(m)
                cols[i].Mask = m.ctypes.data
        self.rows = int(n or 0)
        self.cols = cols
        self.rs = abi.TRowSet()
        self.rs.Columns = C.cast(cols, C.POINTER(abi.TColumn))
        self.rs.ColumnCount = len(self.types)
        self.rs.RowCount = self.rows
        self.rs.RefCount = 1
        if selection is not None:
            s = np.ascontiguousarray(np.asarray(selection, dtype=np.uint8))
            self.keep.append(s)
            self.rs.Selection = s.ctypes.data
        if hashes is not None:
            h = np.ascontiguousarray(np.asarray(hashes, dtype=np.uint64))
            self.keep.append(h)
            self.rs.Hash = h.ctypes.data

    @property
    def ptr(self):
        return C.byref(self.rs)


def read_host_rowset(rs, types):
    """(values, valid) per column of a host TRowSet; 16-byte values as ints."""
    n = rs.RowCount
    out = []
    for i, t in enumerate(types):
        col = rs.Columns[i]
        if t.kind == abi.BOOL:
            raw = np.ctypeslib.as_array(C.cast(col.Data, C.POINTER(C.c_uint8)), shape=(max(1, (n + 7) // 8),)).copy()
            vals = np.unpackbits(raw, bitorder="little", count=n).astype(bool)
        elif n == 0:
            vals = np.zeros(0, dtype=_np_dtype(t))
        else:
            raw = np.ctypeslib.as_array(C.cast(col.Data, C.POINTER(C.c_uint8)), shape=(n * t.width,)).copy()
            vals = raw.view(_np_dtype(t))
        if t.width == 16 and t.kind != abi.BOOL:
            vals = [int(h) * (1 << 64) + int(l) for l, h in zip(vals["lo"].tolist(), vals["hi"].tolist())]
        valid = None
        if col.Mask:
            raw = np.ctypeslib.as_array(C.cast(col.Mask, C.POINTER(C.c_uint8)), shape=(max(1, (n + 7) // 8),)).copy()
            valid = np.unpackbits(raw, bitorder="little", count=n).astype(bool)
        out.append((vals, valid))
    return out


def decimal_to_f64(value: int, scale: int) -> float:
    u = value & ((1 << 128) - 1)
    lo = u & ((1 << 64) - 1)
    hi = u >> 64
    to_s = lambda x: x - (1 << 64) if x >= (1 << 63) else x
    return LIB.oracle_decimal_to_f64(to_s(lo), to_s(hi), scale)


def exact_sum(values) -> float:
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
    return LIB.oracle_exact_sum(a.ctypes.data, len(a))


def kahan_sum(values) -> float:
    a = np.ascontiguousarray(np.asarray(values, dtype=np.float64))
    return LIB.oracle_kahan_sum(a.ctypes.data, len(a))


def filter_rows(spec, rowset: HostRowSet, input_sel=None) -> np.ndarray:
    out = np.zeros(rowset.rows, dtype=np.uint8)
    sel = None
    if input_sel is not None:
        sel = np.ascontiguousarray(np.asarray(input_sel, dtype=np.uint8))
    LIB.oracle_filter(C.byref(spec), rowset.ptr, sel.ctypes.data if sel is not None else None, out.ctypes.data)
    return out


class RowError(RuntimeError):
    """A row of a project failed (qdb_decimal64_overflow)."""


def project_rows(spec, rowset: HostRowSet, out_types, input_sel=None):
    n = max(rowset.rows, 1)
    k = len(out_types)
    bufs = []
    ptrs = (C.c_void_p * (2 * k))()
    for i, t in enumerate(out_types):
        b = np.zeros((n + 7) // 8 + 8 if t.kind == abi.BOOL else n * t.width + 16, dtype=np.uint8)
        bufs.append(b)
        ptrs[i] = b.ctypes.data
        if t.nullable:
            m = np.zeros((n + 7) // 8 + 8, dtype=np.uint8)
            bufs.append(m)
            ptrs[k + i] = m.ctypes.data
    sel = None if input_sel is None else np.ascontiguousarray(np.asarray(input_sel, dtype=np.uint8))
    failed = LIB.oracle_project(C.byref(spec), rowset.ptr, ptrs, sel.ctypes.data if sel is not None else None)
    if failed:
        # The reference's project evaluates every row with row errors
        # deferred, then the selected rows raising (qdb/exec/project_exec.cpp:
        # 71-81): only a failing row that the selection keeps is an error.
        raise RowError(f"value does not fit DECIMAL precision {failed}")
    out = []
    j = 0
    for i, t in enumerate(out_types):
        b = bufs[j]
        j += 1
        if t.kind == abi.BOOL:
            vals = np.unpackbits(b, bitorder="little", count=rowset.rows).astype(bool)
        else:
            vals = b[: rowset.rows * t.width].view(_np_dtype(t))
            if t.width == 16:
                vals = [int(h) * (1 << 64) + int(l) for l, h in zip(vals["lo"].tolist(), vals["hi"].tolist())]
        valid = None
        if t.nullable:
            valid = np.unpackbits(bufs[j], bitorder="little", count=rowset.rows).astype(bool)
            j += 1
        out.append((vals, valid))
    return out


def output_types(spec):
    n = spec.NumKeys + spec.NumAggs
    arr = (abi.QdbValueType * max(1, n))()
    LIB.oracle_agg_output_types(C.byref(spec), arr)
    return arr[:n]


def aggregate(spec, batches: Sequence[HostRowSet], fsum_mode: int = 0, capacity: int = 8):
    """The swiss-table aggregate over the batches; groups in first-seen
    order. fsum_mode: 0 arrival order (the reference), 1 Kahan-Neumaier,
    2 exact."""
    h = LIB.oracle_agg_new(C.byref(spec), capacity, fsum_mode)
    if not h:
        raise ValueError("oracle: unsupported aggregate spec")
    try:
        for b in batches:
            LIB.oracle_agg_update(h, b.ptr)
        out = abi.TRowSet()
        LIB.oracle_agg_finish(h, C.byref(out))
        try:
            types = output_types(spec)

            class _T:
                def __init__(self, t):
                    self.kind, self.width, self.scale, self.nullable = t.Kind, t.Width, t.Scale, bool(t.Nullable)
            return read_host_rowset(out, [_T(t) for t in types])
        finally:
            LIB.oracle_rowset_release(C.byref(out))
    finally:
        LIB.oracle_agg_free(h)


class OracleJoin:
    def __init__(self, spec):
        self.spec = spec
        self.h = LIB.oracle_join_new(C.byref(spec))

    def append(self, b: HostRowSet):
        LIB.oracle_join_append(self.h, b.ptr)

    def build(self):
        LIB.oracle_join_build(self.h)

    def probe(self, b: HostRowSet) -> np.ndarray:
        n = LIB.oracle_join_probe(self.h, b.ptr)
        if n == 0:
            return np.zeros((0, 2), dtype=np.int64)
        p = np.ctypeslib.as_array(LIB.oracle_join_pairs(self.h), shape=(2 * n,)).copy()
        return p.reshape(n, 2)

    def keys(self) -> int:
        return LIB.oracle_join_keys(self.h)

    def rows(self) -> int:
        return LIB.oracle_join_rows(self.h)

    def __del__(self):
        try:
            LIB.oracle_join_free(self.h)
        except Exception:
            pass


def gen_lineitem(seed: int, row0: int, rows: int):
    cols = [np.empty(rows, np.int8), np.empty(rows, np.int8), np.empty(rows, np.int32)] + [np.empty(rows, np.int64) for _ in range(4)]
    LIB.oracle_gen_lineitem(seed, row0, rows, *[c.ctypes.data for c in cols])
    return cols


def gen_join_build(seed: int, row0: int, rows: int, build_rows: int):
    key, seg = np.empty(rows, np.int64), np.empty(rows, np.int32)
    LIB.oracle_gen_join_build(seed, row0, rows, build_rows, key.ctypes.data, seg.ctypes.data)
    return key, seg


def gen_join_probe(seed: int, row0: int, rows: int, build_rows: int):
    fk, val = np.empty(rows, np.int64), np.empty(rows, np.int64)
    LIB.oracle_gen_join_probe(seed, row0, rows, build_rows, fk.ctypes.data, val.ctypes.data)
    return fk, val


def zipf_cdf(n: int) -> np.ndarray:
    cdf = np.empty(n, np.float64)
    LIB.oracle_zipf_cdf(cdf.ctypes.data, n)
    return cdf


def gen_groups(seed: int, row0: int, rows: int, cdf: np.ndarray):
    key, vi, vf = np.empty(rows, np.int64), np.empty(rows, np.int64), np.empty(rows, np.float64)
    LIB.oracle_gen_groups(seed, row0, rows, cdf.ctypes.data, len(cdf), key.ctypes.data, vi.ctypes.data, vf.ctypes.data)
    return key, vi, vf


class OracleRuntimeFilter:
    """The reference's runtime join filter for one producer (oracle.c)."""

    def __init__(self, max_keys: int = 1 << 14, bloom_bytes: int = 2 << 20):
        self.f = LIB.oracle_rf_new(max_keys, bloom_bytes)

    def build(self, hashes, selection=None):
        h = np.ascontiguousarray(np.asarray(hashes, dtype=np.uint64))
        s = None if selection is None else np.ascontiguousarray(np.asarray(selection, dtype=np.uint8))
        LIB.oracle_rf_build(self.f, h.ctypes.data, s.ctypes.data if s is not None else None, len(h))

    def publish(self) -> int:
        return LIB.oracle_rf_publish(self.f)

    def probe(self, hashes, selection=None):
        h = np.ascontiguousarray(np.asarray(hashes, dtype=np.uint64))
        s = None if selection is None else np.ascontiguousarray(np.asarray(selection, dtype=np.uint8))
        out = np.zeros(len(h), np.uint8)
        kept = LIB.oracle_rf_probe(self.f, h.ctypes.data, s.ctypes.data if s is not None else None, out.ctypes.data, len(h))
        return out, int(kept)

    def __del__(self):
        try:
            LIB.oracle_rf_free(self.f)
        except Exception:
            pass


def key_hashes(rowset: HostRowSet, types_c, key_cols: Sequence[int]) -> np.ndarray:
    """rh_hash of the key columns of every row (oracle_key_hash); types_c:
    the columns' QdbValueType structs."""
    tt = (abi.QdbValueType * max(1, len(types_c)))(*[abi.QdbValueType.from_buffer_copy(t) for t in types_c])
    kk = (C.c_int32 * max(1, len(key_cols)))(*key_cols)
    n = rowset.rs.RowCount
    return np.array([LIB.oracle_key_hash(rowset.rs.Columns, tt, kk, len(key_cols), i) for i in range(n)], dtype=np.uint64)


def shuffle_partition(hashes: np.ndarray, selection: Optional[np.ndarray], outputs: int,
                      routes: Optional[Sequence[int]] = None):
    """The hash shuffle's routing of one batch: every selected row, in input
    order, goes to output hash % outputs, or with a routing table to
    routes[hash % len(routes)], where an entry outside [0, outputs) is an
    error (THashShuffleTask::OutputOf, qdb/scheduler/tasks/
    runtime_adapter.cpp:998-1010, over the row loop of :1100-1105).
    Returns (row_index, offsets): output o's rows at
    row_index[offsets[o]:offsets[o + 1]]."""
    per = [[] for _ in range(outputs)]
    for row in range(len(hashes)):
        if selection is not None and not selection[row]:
            continue
        h = int(hashes[row])
        if routes is None:
            out = h % outputs
        else:
            part = h % len(routes)
            out = int(routes[part])
            if out < 0 or out >= outputs:
                raise ValueError(f"hash shuffle got a row of partition {part}, which it routes nowhere")
        per[out].append(row)
    offsets = np.zeros(outputs + 1, dtype=np.int64)
    for o in range(outputs):
        offsets[o + 1] = offsets[o] + len(per[o])
    rows = np.array([r for p in per for r in p], dtype=np.uint32)
    return rows, offsets


def partial_states(spec, rowset: HostRowSet):
    """The single-row partial states of a batch's selected rows
    (BuildAggregatePassThroughOperator, qdb/exec/planner_helpers.cpp:211-260):
    the keys, 1 for COUNT(*), 0 for a NULL argument of COUNT(x) and 1
    otherwise, the argument of SUM/MIN/MAX as it is (the state type widens
    it). Returns per output column (values, valid) over the selected rows."""
    sel = np.ones(rowset.rs.RowCount, dtype=bool) if rowset.selection is None else np.asarray(rowset.selection, bool)
    out = []
    for k in range(spec.NumKeys):
        c = spec.KeyColumns[k]
        v = rowset.valid[c]
        out.append((np.asarray(rowset.columns[c])[sel] if not isinstance(rowset.columns[c], list)
                    else [x for x, s in zip(rowset.columns[c], sel) if s], None if v is None else np.asarray(v)[sel]))
    for a in range(spec.NumAggs):
        fn = spec.Aggs[a]
        if fn.Func == abi.AGG_COUNT_STAR:
            out.append((np.ones(int(sel.sum()), dtype=np.int64), None))
        elif fn.Func == abi.AGG_COUNT:
            v = rowset.valid[fn.Column]
            ones = np.ones(rowset.rs.RowCount, dtype=np.int64) if v is None else np.asarray(v).astype(np.int64)
            out.append((ones[sel], None))
        else:
            c = fn.Column
            v = rowset.valid[c]
            col = rowset.columns[c]
            vals = np.asarray(col)[sel] if not isinstance(col, list) else [x for x, s in zip(col, sel) if s]
            out.append((vals, None if v is None else np.asarray(v)[sel]))
    return out


class PartialAggregate:
    """The partial aggregate's policy (TPartialAggregateProcessor::Add,
    qdb/exec/aggregate_exec.cpp:180-216): after each batch the table's group
    count decides whether the table is emitted (max_groups) or whether the
    lane stops hashing (sample_rows rows at fewer than min_reduction rows per
    group, emitted groups included). The table is the oracle aggregate of
    the batches since the last emission. outputs lists ("table", result) and
    ("states", batch) in order."""

    def __init__(self, spec, sample_rows=128 * 1024, min_reduction=8, max_groups=1 << 20, pass_through=True):
        self.spec = spec
        self.sample_rows = sample_rows
        self.min_reduction = min_reduction
        self.max_groups = max_groups
        self.pass_through = pass_through
        self.table = []
        self.hashed_rows = 0
        self.emitted_groups = 0
        self.passing = False
        self.outputs = []

    def _groups(self):
        if not self.table:
            return 0
        return len(aggregate(self.spec, self.table)[0][0]) if self.spec.NumKeys else 1

    def _flush(self):
        if self.table and self._groups() > 0:
            self.outputs.append(("table", aggregate(self.spec, self.table)))
        self.table = []

    def add(self, batch: HostRowSet):
        if self.passing:
            self.outputs.append(("states", batch))
            return
        self.hashed_rows += batch.rs.RowCount
        self.table.append(batch)
        groups = self._groups()
        if (self.pass_through and self.hashed_rows >= self.sample_rows
                and self.hashed_rows < self.min_reduction * (self.emitted_groups + groups)):
            self.passing = True
            self._flush()
            return
        if groups >= self.max_groups:
            self.emitted_groups += groups
            self._flush()


def _radix_key_values(t, vals):
    """qumir_radix_key (qdb/kernel/sort/radix.oz:4-60) of every value, as
    Python ints; with the key's bit count."""
    out = []
    if t.kind == abi.BOOL:
        return [1 if v else 0 for v in vals], 8
    if t.kind == abi.FLOAT:
        w = 8 if t.width == 8 else 4
        for v in vals:
            b = int(np.array([v], dtype=np.float64 if w == 8 else np.float32).view(np.uint64 if w == 8 else np.uint32)[0])
            top = 1 << (8 * w - 1)
            mag = b & (top - 1)
            inf = 0x7FF0000000000000 if w == 8 else 0x7F800000
            if mag > inf:
                out.append((1 << (8 * w)) - 1)
                continue
            if mag == 0:
                b = 0
            out.append(b ^ (((1 << (8 * w)) - 1) if b >> (8 * w - 1) else top))
        return out, 8 * w
    bits = 8 * t.width
    for v in vals:
        v = int(v)
        if t.kind == abi.UNSIGNED:
            out.append(v)
        else:
            out.append((v + (1 << (bits - 1))) & ((1 << bits) - 1))
    return out, bits


def sort_rows(types, columns, valid, selection, keys):
    """ORDER BY as the reference's radix sort of row ids: the selected rows,
    then for each key from last to first, for each 8-bit digit from the
    lowest, a stable counting sort by the digit (255 - digit descending);
    a nullable key's NULL rows in a bucket before (NULLS FIRST) or after
    every digit (radix_count_sort_indices / _nullable,
    qdb/kernel/sort/radix.oz:62-138, radix_nullable.oz:5-101; the default
    placement of qdb/plan/ops/sort_key.h:21-26). keys: (column, desc,
    nulls) with nulls 0 default, 1 first, 2 last. Returns row indexes."""
    n = len(columns[0]) if columns else 0
    order = [i for i in range(n) if selection is None or selection[i]]
    for col, desc, nulls in reversed([tuple(k) for k in keys]):
        t = types[col]
        rk, bits = _radix_key_values(t, columns[col])
        v = valid[col] if valid is not None else None
        first = desc if nulls == 0 else nulls == 1
        for digit in range(0, bits, 8):
            def bucket(row):
                if v is not None and not v[row]:
                    return 0 if first else 256
                d = (rk[row] >> digit) & 255
                if desc:
                    d = 255 - d
                return d + 1 if (v is not None and first) else d
            order = sorted(order, key=bucket)  # Python's sort is stable
    return order
