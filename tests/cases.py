"""Backend-neutral parity cases: the reference's golden queries for the path,
restated over the C-ABI specs, and two backends that run them: the CPU
oracle and the HIP library. The oracle tests pin the oracle against the
goldens; the GPU tests pin the HIP path against the goldens and the oracle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from helpers import (DECIMAL, FLOAT, SIGNED, T, golden_rows, load_fixtures, np_column, result_rows, sort_key,
                     table_columns)

# Op-codes of include/qdb_gpu.h (mirrored so that the oracle tests do not
# load the HIP library).
AGG_COUNT_STAR, AGG_COUNT, AGG_SUM, AGG_MIN, AGG_MAX = range(5)
JOIN_INNER, JOIN_LEFT_OUTER, JOIN_RIGHT_OUTER = range(3)


@dataclass
class Batch:
    types: list
    columns: list
    valid: Optional[list] = None
    selection: Optional[np.ndarray] = None


def _vt(t):
    """helpers.T -> qumirdb_amd.host.ValueType-like object used by the spec
    builders (kind, width, scale, nullable, precision)."""
    from oracle.oracle import abi  # noqa: F401  (struct defs only)
    return t


# ---------------------------------------------------------------------------
# Spec construction without importing the HIP library
# ---------------------------------------------------------------------------

def _abi():
    from oracle.oracle import abi
    return abi


def value_type_c(t):
    return _abi().QdbValueType(t.kind, t.width, t.scale, bool(t.nullable), t.precision if t.kind == DECIMAL else 0)


def schema_c(types):
    abi = _abi()
    s = abi.QdbColumnSchema()
    s.NumColumns = len(types)
    for i, t in enumerate(types):
        s.Types[i] = value_type_c(t)
    return s


class Prog:
    """Minimal program builder over helpers.T types (mirrors host.Program)."""

    def __init__(self, schema):
        self.schema = list(schema)
        self.insns = []
        self.types = []

    def _emit(self, op, cls, typ, a=0, b=0, column=0, lo=0, hi=0):
        abi = _abi()
        ins = abi.QdbExprInsn()
        dst = len(self.types)
        assert dst < abi.MAX_REGS
        ins.Op, ins.Dst, ins.A, ins.B, ins.Class, ins.Column, ins.ImmLo, ins.ImmHi = op, dst, a, b, cls, column, lo, hi
        self.insns.append(ins)
        self.types.append(typ)
        return dst

    def col(self, i):
        abi = _abi()
        t = self.schema[i]
        cls = abi.REG_F64 if t.kind == FLOAT else abi.REG_INT
        return self._emit(abi.EXPR_COLUMN, cls, t, column=i)

    def const(self, v, typ):
        abi = _abi()
        u = int(v) & ((1 << 128) - 1)
        s = lambda x: x - (1 << 64) if x >= (1 << 63) else x
        return self._emit(abi.EXPR_CONST_INT, abi.REG_INT, typ, lo=s(u & ((1 << 64) - 1)), hi=s(u >> 64))

    def const_f64(self, v):
        abi = _abi()
        bits = int(np.array([v], dtype=np.float64).view(np.int64)[0])
        return self._emit(abi.EXPR_CONST_F64, abi.REG_F64, T(FLOAT, 8), lo=bits)

    def op(self, op, a, b, typ, cls=None):
        abi = _abi()
        return self._emit(op, abi.REG_INT if cls is None else cls, typ, a=a, b=b)

    def scale_up(self, a, k, typ):
        abi = _abi()
        return self._emit(abi.EXPR_SCALE_UP, abi.REG_INT, typ, a=a, lo=k)

    def cmp(self, op, a, b):
        abi = _abi()
        return self._emit(abi.EXPR_CMP, abi.REG_BOOL, T(3, 1, nullable=True), a=a, b=b, lo=op)

    def build(self, result):
        abi = _abi()
        p = abi.QdbExprProgram()
        p.NumInsns = len(self.insns)
        p.Result = result
        for k, ins in enumerate(self.insns):
            p.Insns[k] = ins
        return p


@dataclass
class AggCase:
    types: list
    keys: list
    aggs: list
    global_agg: bool = False
    computed: list = field(default_factory=list)  # [(Prog, result, T)]
    predicate: Optional[tuple] = None             # (Prog, result)
    empty_groups: list = field(default_factory=list)

    def to_c(self):
        abi = _abi()
        s = abi.QdbAggregateSpec()
        s.Input = schema_c(self.types)
        s.NumKeys = len(self.keys)
        for i, k in enumerate(self.keys):
            s.KeyColumns[i] = k
        s.NumAggs = len(self.aggs)
        for i, (f, c) in enumerate(self.aggs):
            s.Aggs[i] = abi.QdbAggregateFn(f, c)
        s.NumEmptyInputGroups = len(self.empty_groups)
        for i, g in enumerate(self.empty_groups):
            s.EmptyInputGroups[i] = g
        if self.predicate is not None:
            s.HasPredicate = 1
            s.Predicate = self.predicate[0].build(self.predicate[1])
        s.NumComputed = len(self.computed)
        for k, (p, r, t) in enumerate(self.computed):
            s.ComputedTypes[k] = value_type_c(t)
            s.Computed[k] = p.build(r)
        return s

    def all_types(self):
        return list(self.types) + [t for _, _, t in self.computed]

    def out_types(self):
        types = self.all_types()
        out = [types[k] for k in self.keys]
        for f, c in self.aggs:
            if f in (AGG_COUNT_STAR, AGG_COUNT):
                out.append(T(SIGNED, 8))
                continue
            t = types[c]
            nullable = t.nullable or self.global_agg or bool(self.empty_groups)
            if t.kind == FLOAT:
                out.append(T(FLOAT, 8, nullable=nullable))
            elif t.kind == DECIMAL:
                out.append(T(DECIMAL, 16 if f == AGG_SUM else t.width, t.scale, nullable))
            else:
                out.append(T(SIGNED, 8, nullable=nullable))
        return out


def lower_global(case: AggCase, batches):
    """A global aggregate (no GROUP BY) as the reference plans it
    (qdb/plan/build.cpp:2461-2524): the input gains the constant i64 column
    `__group__` = 1 after its own columns, the aggregate groups by it and seeds
    the group 1 that exists over an empty input. Returns the lowered case,
    the extended batches and whether the leading key column is to be dropped
    from the result."""
    if not case.global_agg:
        return case, batches, False
    nin = len(case.types)
    aggs = [(f, c + 1 if c >= nin else c) for f, c in case.aggs]
    lowered = AggCase(list(case.types) + [T(SIGNED, 8)], [nin], aggs, computed=case.computed, predicate=case.predicate,
                      empty_groups=[1])
    out = []
    for b in batches:
        n = len(b.columns[0]) if b.columns else 0
        valid = None if b.valid is None else list(b.valid) + [None]
        out.append(Batch(list(b.types) + [T(SIGNED, 8)], list(b.columns) + [np.ones(n, np.int64)], valid, b.selection))
    return lowered, out, True


@dataclass
class JoinCase:
    build_types: list
    probe_types: list
    build_keys: list
    probe_keys: list
    join_type: int = JOIN_INNER
    probe_is_left: bool = True
    fused: list = field(default_factory=list)
    table_mode: int = 0     # 0: the build picks (dense direct table or hash), 1: hash table

    def to_c(self):
        abi = _abi()
        s = abi.QdbRowJoinSpec()
        s.Build = schema_c(self.build_types)
        s.Probe = schema_c(self.probe_types)
        s.NumKeys = len(self.build_keys)
        for i, (b, p) in enumerate(zip(self.build_keys, self.probe_keys)):
            s.BuildKeyColumns[i] = b
            s.ProbeKeyColumns[i] = p
        s.JoinType = self.join_type
        s.ProbeIsLeft = int(self.probe_is_left)
        s.NumFused = len(self.fused)
        for i, (f, c) in enumerate(self.fused):
            s.Fused[i] = abi.QdbAggregateFn(f, c)
        s.TableMode = self.table_mode
        return s

    def out_types(self):
        outer = self.join_type != JOIN_INNER
        build = [T(t.kind, t.width, t.scale, t.nullable or outer, t.precision) for t in self.build_types]
        return (list(self.probe_types) + build) if self.probe_is_left else (build + list(self.probe_types))


# ---------------------------------------------------------------------------
# Backends
# ---------------------------------------------------------------------------

class OracleBackend:
    name = "oracle"

    def __init__(self, fsum_mode=2):
        from oracle import oracle as O
        self.O = O
        self.fsum_mode = fsum_mode

    def host(self, b: Batch):
        return self.O.HostRowSet(b.types, [np_column(t, c) for t, c in zip(b.types, b.columns)], b.valid, b.selection)

    def aggregate(self, case: AggCase, batches):
        case, batches, strip = lower_global(case, batches)
        res = self.O.aggregate(case.to_c(), [self.host(b) for b in batches], fsum_mode=self.fsum_mode)
        return res[1:] if strip else res

    def join(self, case: JoinCase, build_batches, probe: Batch):
        """(pairs, output rows) of a join; output rows gathered on the host."""
        j = self.O.OracleJoin(case.to_c())
        for b in build_batches:
            j.append(self.host(b))
        j.build()
        pairs = j.probe(self.host(probe))
        return pairs, gather_rows(case, build_batches, probe, pairs)

    def filter(self, types, prog, res, batch: Batch):
        abi = _abi()
        spec = abi.QdbFilterSpec()
        spec.Input = schema_c(types)
        spec.Predicate = prog.build(res)
        return self.O.filter_rows(spec, self.host(batch), batch.selection)

    def project(self, types, outputs, batch: Batch):
        abi = _abi()
        spec = abi.QdbProjectSpec()
        spec.Input = schema_c(types)
        spec.NumOutputs = len(outputs)
        for k, (p, r, t) in enumerate(outputs):
            spec.OutputTypes[k] = value_type_c(t)
            spec.Outputs[k] = p.build(r)
        hb = self.host(Batch(batch.types, batch.columns, batch.valid))
        return self.O.project_rows(spec, hb, [t for _, _, t in outputs], batch.selection)


class GpuBackend:
    name = "gpu"

    def __init__(self, init=True):
        import qumirdb_amd as q
        from qumirdb_amd import host as H
        if init:
            q.init_device(0)
        self.q, self.H = q, H

    def cast(self, s):
        """The same struct as the HIP library's binding class (the oracle loads
        its own copy of the ABI module)."""
        return getattr(self.q.abi, type(s).__name__).from_buffer_copy(s)

    def dev(self, b: Batch):
        H = self.H
        types = [H.ValueType(t.kind, t.width, t.scale, t.nullable, t.precision) for t in b.types]
        cols = [np_column(t, c) for t, c in zip(b.types, b.columns)]
        return H.DeviceRowSet(types, cols, b.valid, b.selection, rows=len(b.columns[0]) if b.columns else 0)

    def aggregate(self, case: AggCase, batches, capacity=8):
        case, batches, strip = lower_global(case, batches)
        res = self._aggregate(case, batches, capacity)
        return res[1:] if strip else res

    def _aggregate(self, case: AggCase, batches, capacity=8):
        H = self.H
        kern = H.AggregateKernels.__new__(H.AggregateKernels)
        kern.spec = None
        kern.c_spec = self.cast(case.to_c())
        kern.k = self.q.abi.QdbAggregateKernels()
        self.q.check(self.q.lib().qdb_gpu_compile_aggregate(C.byref(kern.c_spec), C.byref(kern.k)), "compile_aggregate")
        kern.out_types = [H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in kern.k.OutputTypes[: kern.k.NumOutputColumns]]
        proc = H.TAggregateProcessor(kern, initial_capacity=capacity)
        devs = [self.dev(b) for b in batches]
        for d in devs:
            proc.add(d)
        out = proc.finish()
        proc.close()
        self.q.lib().qdb_gpu_release_slot(self.q.abi.SLOT_AGGREGATE, kern.k.Slot)
        return out

    def join_processor(self, case: JoinCase):
        H = self.H
        kern = H.RowJoinKernels.__new__(H.RowJoinKernels)
        kern.spec = case
        kern.c_spec = self.cast(case.to_c())
        kern.k = self.q.abi.QdbRowJoinKernels()
        self.q.check(self.q.lib().qdb_gpu_compile_row_join(C.byref(kern.c_spec), C.byref(kern.k)), "compile_row_join")
        kern.out_types = [H.ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in kern.k.OutputTypes[: kern.k.NumOutputColumns]]
        return kern, H.TRowJoinProcessor(kern, build_is_left=not case.probe_is_left)

    def join(self, case: JoinCase, build_batches, probe: Batch):
        kern, proc = self.join_processor(case)
        for b in build_batches:
            proc.append(self.dev(b))
        proc.build()
        pd = self.dev(probe)
        pairs = proc.probe_pairs(pd)
        outs = proc.probe(pd)
        rows = []
        for cols in outs:
            rows.extend(result_rows(cols))
        proc.close()
        self.q.lib().qdb_gpu_release_slot(self.q.abi.SLOT_JOIN, kern.k.Slot)
        return pairs, rows

    def filter(self, types, prog, res, batch: Batch):
        abi = self.q.abi
        spec = abi.QdbFilterSpec()
        spec.Input = abi.QdbColumnSchema.from_buffer_copy(schema_c(types))
        spec.Predicate = abi.QdbExprProgram.from_buffer_copy(prog.build(res))
        fn = C.c_void_p()
        slot = self.q.check(self.q.lib().qdb_gpu_compile_filter(C.byref(spec), C.byref(fn)), "compile_filter")
        f = abi.FILTER_FN(fn.value)
        d = self.dev(Batch(batch.types, batch.columns, batch.valid))
        out = self.H.DeviceBuffer(max(d.rows, 1))
        view = abi.TRowSet.from_buffer_copy(d.rs)
        view.Selection = out.ptr
        insel = self.H.DeviceBuffer.from_numpy(np.asarray(batch.selection, np.uint8)) if batch.selection is not None else None
        f(C.byref(view), None, None, insel.ptr if insel is not None else None)
        self.q.lib().qdb_gpu_release_slot(abi.SLOT_FILTER, slot)
        return out.to_numpy(np.uint8, d.rows)

    def project(self, types, outputs, batch: Batch):
        abi = self.q.abi
        H = self.H
        spec = abi.QdbProjectSpec()
        spec.Input = abi.QdbColumnSchema.from_buffer_copy(schema_c(types))
        spec.NumOutputs = len(outputs)
        for k, (p, r, t) in enumerate(outputs):
            spec.OutputTypes[k] = abi.QdbValueType.from_buffer_copy(value_type_c(t))
            spec.Outputs[k] = abi.QdbExprProgram.from_buffer_copy(p.build(r))
        fn = C.c_void_p()
        slot = self.q.check(self.q.lib().qdb_gpu_compile_project(C.byref(spec), C.byref(fn)), "compile_project")
        f = abi.PROJECT_FN(fn.value)
        d = self.dev(Batch(batch.types, batch.columns, batch.valid))
        n = max(d.rows, 1)
        k = len(outputs)
        bufs, ptrs = [], (C.c_void_p * (2 * k))()
        masks = {}
        for i, (_, _, t) in enumerate(outputs):
            b = H.DeviceBuffer((n + 7) // 8 + 8 if t.kind == 3 else n * t.width)
            b.zero()
            bufs.append(b)
            ptrs[i] = b.ptr
            if t.nullable:
                m = H.DeviceBuffer((n + 7) // 8 + 8)
                m.zero()
                masks[i] = m
                ptrs[k + i] = m.ptr
        insel = H.DeviceBuffer.from_numpy(np.asarray(batch.selection, np.uint8)) if batch.selection is not None else None
        lib = self.q.lib()
        try:
            # The reference's project operator (qdb/exec/project_exec.cpp:
            # 71-81): every row with row errors deferred; after a row error,
            # the selected rows again, now raising.
            f(C.byref(d.rs), ptrs, None, None, None)
            if lib.qdb_gpu_take_row_error():
                f(C.byref(d.rs), ptrs, None, None, insel.ptr if insel is not None else None)
                if lib.qdb_gpu_take_row_error():
                    raise RuntimeError(self.q.last_error())
            elif insel is not None:
                f(C.byref(d.rs), ptrs, None, None, insel.ptr)
        finally:
            lib.qdb_gpu_release_slot(abi.SLOT_PROJECT, slot)
        out = []
        for i, (_, _, t) in enumerate(outputs):
            if t.kind == 3:
                vals = H.unpack_bitmap(bufs[i].to_numpy(np.uint8, (d.rows + 7) // 8), d.rows)
            else:
                vt = H.ValueType(t.kind, t.width, t.scale, t.nullable)
                vals = bufs[i].to_numpy(vt.np_dtype, d.rows)
                if t.width == 16:
                    vals = H.i128_values(vals)
            valid = H.unpack_bitmap(masks[i].to_numpy(np.uint8, (d.rows + 7) // 8), d.rows) if i in masks else None
            out.append((vals, valid))
        return out


class ProtocolBackend(GpuBackend):
    """The HIP library's aggregate kernels driven by the reference's own
    protocol (tests/native/ref_protocol.c: a task heap with owner headers,
    TAggregateProcessor's Init / Update / Size == finalized check, and the
    AdoptKernelOwnedRowSet / DestroyKernelOwnedRowSet sequence), with the
    outputs in host memory from that heap (OutputOnHost = 1)."""

    def __init__(self):
        import os
        super().__init__()
        self.h = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "native", "libref_protocol.so"))
        self.h.harness_install.argtypes = [C.c_void_p]
        self.h.harness_counters.argtypes = [C.c_void_p]
        self.h.harness_aggregate.restype = C.c_int64
        self.h.harness_aggregate.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]
        self.h.harness_install(C.cast(self.q.lib().qdb_gpu_set_host_heap, C.c_void_p))

    def counters(self):
        out = (C.c_int64 * 3)()
        self.h.harness_counters(out)
        return tuple(out)

    def uninstall(self):
        """Back to the library's default heap, once no block of this one is
        alive (a block must go back to the heap it came from)."""
        if self.counters()[0] != 0:
            return
        abi = self.q.abi
        self.q.lib().qdb_gpu_set_host_heap(abi.HEAP_ALLOC_FN(0), abi.HEAP_FREE_FN(0), abi.HEAP_DETACH_FN(0))

    def _aggregate(self, case: AggCase, batches, capacity=8):
        from oracle.oracle import read_host_rowset
        H, abi, lib = self.H, self.q.abi, self.q.lib()
        spec = self.cast(case.to_c())
        spec.OutputOnHost = 1
        k = abi.QdbAggregateKernels()
        self.q.check(lib.qdb_gpu_compile_aggregate(C.byref(spec), C.byref(k)), "compile_aggregate")
        types = [T(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in k.OutputTypes[: k.NumOutputColumns]]
        devs = [self.dev(b) for b in batches]
        arr = (C.POINTER(abi.TRowSet) * max(1, len(devs)))(*[C.pointer(d.rs) for d in devs])
        before = self.counters()
        out = abi.TRowSet()
        rows = self.h.harness_aggregate(C.byref(k), arr, len(devs), capacity, C.byref(out))
        lib.qdb_gpu_release_slot(abi.SLOT_AGGREGATE, k.Slot)
        assert rows != -2, "agg_finish_rowset returned a row count other than the table's Size"
        assert rows >= 0, f"aggregate failed: {self.q.last_error()}"
        adopted = self.counters()
        # Adopted: every owner (and the list) detached from the task heap,
        # each one a block of that heap.
        assert adopted[1] == before[1] and adopted[2] == 0, (before, adopted)
        assert out.Destroy, "the adoption installs the destroy callback"
        res = read_host_rowset(out, types)
        lib.qdb_gpu_rowset_release(C.byref(out))
        after = self.counters()
        assert after[0] == before[0] and after[2] == 0, f"blocks left or foreign frees: {before} -> {after}"
        return res


def gather_rows(case: JoinCase, build_batches, probe: Batch, pairs):
    """Output rows of a join from its pairs, gathered on the host."""
    bcols = [[] for _ in case.build_types]
    bvalid = [[] for _ in case.build_types]
    for b in build_batches:
        for c in range(len(case.build_types)):
            bcols[c].extend(list(b.columns[c]))
            v = b.valid[c] if b.valid is not None and b.valid[c] is not None else np.ones(len(b.columns[c]), bool)
            bvalid[c].extend(list(v))
    rows = []
    for a, b in pairs:
        stream, build = (a, b) if case.probe_is_left else (b, a)
        pr = int(stream) + (1 << 32)
        prow = []
        for c in range(len(case.probe_types)):
            ok = probe.valid is None or probe.valid[c] is None or probe.valid[c][pr]
            prow.append(int(probe.columns[c][pr]) if ok else None)
        brow = []
        for c in range(len(case.build_types)):
            if build < 0 or not bvalid[c][build]:
                brow.append(None)
            else:
                brow.append(int(bcols[c][build]))
        rows.append(tuple(prow + brow) if case.probe_is_left else tuple(brow + prow))
    return rows


# ---------------------------------------------------------------------------
# Golden cases
# ---------------------------------------------------------------------------

FIX = load_fixtures()


def table_batch(name, selection=None):
    names, types, values, valid = table_columns(FIX["tables"][name])
    return names, Batch(types, values, valid, selection)


def backend_avg(backend, sums, counts, scale, kind=DECIMAL, width=16):
    """AVG over finished groups as the reference lowers it
    (qdb/plan/sql_lowering.cpp:799-814): SUM / COUNT, both converted to DOUBLE
    (the decimal correctly rounded, qdb_decimal_to_f64) and divided, as a
    projection that the backend runs (on the GPU: the project kernel)."""
    E = _abi()
    st = T(kind, width, scale, True, 38 if width == 16 else 18)
    ct = T(SIGNED, 8)
    types = [st, ct]
    p = Prog(types)
    num = p._emit(E.EXPR_TO_F64, E.REG_F64, T(FLOAT, 8, nullable=True), a=p.col(0), lo=scale)
    den = p._emit(E.EXPR_TO_F64, E.REG_F64, T(FLOAT, 8), a=p.col(1), lo=0)
    q = p.op(E.EXPR_DIV, num, den, T(FLOAT, 8, nullable=True), cls=E.REG_F64)
    out_t = T(FLOAT, 8, nullable=True)
    valid = np.array([s_ is not None and bool(c) for s_, c in zip(sums, counts)])
    svals = [int(s_ or 0) for s_ in sums]
    if width != 16:
        svals = np.array(svals, dtype=np.int64)
    batch = Batch(types, [svals, np.array([int(c) for c in counts], np.int64)], [valid, None])
    (vals, ok), = backend.project(types, [(p, q, out_t)], batch)
    return [float(v) if ok[i] else None for i, v in enumerate(vals)]


def golden_decimal_aggregates(backend):
    """decimal_parquet/aggregates: COUNT(*), COUNT/SUM/AVG/MIN/MAX of d2 and
    d10 over the wide table, a global aggregate (AVG = SUM / COUNT)."""
    names, b = table_batch("decimal_parquet.wide")
    D2, D10 = names.index("d2"), names.index("d10")
    case = AggCase(b.types, [], [(AGG_COUNT_STAR, -1), (AGG_COUNT, D2), (AGG_SUM, D2), (AGG_MIN, D2), (AGG_MAX, D2),
                                 (AGG_COUNT, D10), (AGG_SUM, D10), (AGG_MIN, D10), (AGG_MAX, D10)], global_agg=True)
    r = result_rows(backend.aggregate(case, [b]))
    assert len(r) == 1
    n, c2, s2, lo2, hi2, c10, s10, lo10, hi10 = r[0]
    (a2,) = backend_avg(backend, [s2], [c2], 2)
    (a10,) = backend_avg(backend, [s10], [c10], 10)
    got = [(n, c2, s2, a2, lo2, hi2, c10, s10, a10, lo10, hi10)]
    return got, golden_rows(FIX["goldens"]["decimal_parquet/aggregates"])


def golden_decimal_group_by(backend):
    """decimal_parquet/group_by: GROUP BY k, d2 (nullable i32 and 16-byte
    decimal keys) with COUNT(*), SUM/MIN/MAX(d10)."""
    names, b = table_batch("decimal_parquet.wide")
    K, D2, D10 = names.index("k"), names.index("d2"), names.index("d10")
    case = AggCase(b.types, [K, D2], [(AGG_COUNT_STAR, -1), (AGG_SUM, D10), (AGG_MIN, D10), (AGG_MAX, D10)])
    got = sorted(result_rows(backend.aggregate(case, [b])), key=sort_key)
    return got, sorted(golden_rows(FIX["goldens"]["decimal_parquet/group_by"]), key=sort_key)


def _t_columns():
    names, types, values, valid = table_columns(FIX["tables"]["decimal_result.t"])
    return names, types, values, valid


def golden_multiply_sum(backend):
    """decimal_result/multiply_sum: GROUP BY id % 2 of SUM(a2 * a2),
    SUM(a2 * (1 - a2)), AVG(a2 * b0), COUNT(*), where a2 = a as DECIMAL(10,2)
    and b0 = b as DECIMAL(19,0). id % 2 is computed on the host."""
    names, types, values, valid = _t_columns()
    ID, A, B = names.index("id"), names.index("a"), names.index("b")
    g = [int(v) % 2 for v in values[ID]]
    types2 = types + [T(SIGNED, 4)]
    values2 = values + [g]
    valid2 = valid + [None]
    G = len(types)
    a2 = T(DECIMAL, 8, 2, True, 10)
    b0 = T(DECIMAL, 8, 0, True, 19)

    def cast_a(p):
        return p.scale_up(p.col(A), 2, a2)

    p1 = Prog(types2)
    x = cast_a(p1)
    y = cast_a(p1)
    sq = p1.op(_abi().EXPR_MUL, x, y, T(DECIMAL, 16, 4, True, 20))
    p2 = Prog(types2)
    x = cast_a(p2)
    one = p2.const(100, T(DECIMAL, 8, 2, False, 3))
    y = cast_a(p2)
    d = p2.op(_abi().EXPR_SUB, one, y, T(DECIMAL, 8, 2, True, 11))
    disc = p2.op(_abi().EXPR_MUL, x, d, T(DECIMAL, 16, 4, True, 21))
    p3 = Prog(types2)
    x = cast_a(p3)
    bb = p3.col(B)
    ab = p3.op(_abi().EXPR_MUL, x, bb, T(DECIMAL, 16, 2, True, 29))
    computed = [(p1, sq, T(DECIMAL, 16, 4, True, 20)), (p2, disc, T(DECIMAL, 16, 4, True, 21)),
                (p3, ab, T(DECIMAL, 16, 2, True, 29))]
    nin = len(types2)
    case = AggCase(types2, [G], [(AGG_SUM, nin), (AGG_SUM, nin + 1), (AGG_SUM, nin + 2), (AGG_COUNT, nin + 2),
                                (AGG_COUNT_STAR, -1)], computed=computed)
    r = result_rows(backend.aggregate(case, [Batch(types2, values2, valid2)]))
    avgs = backend_avg(backend, [row[3] for row in r], [row[4] for row in r], 2)
    got = sorted([(gk, s, dsc, a, n) for (gk, s, dsc, sab, cab, n), a in zip(r, avgs)], key=sort_key)
    return got, golden_rows(FIX["goldens"]["decimal_result/multiply_sum"])


def golden_storage_sum(backend):
    """decimal_result/storage_sum over t x t (25 rows): SUM of the constant
    999999999999999999 as DECIMAL(18,0), MIN(t1.a as DECIMAL(18,2)),
    MAX(t2.a as DECIMAL(18,2)), COUNT(*). The cross product is built on the
    host; the decimal casts scale the i32 a by 100."""
    names, types, values, valid = _t_columns()
    A = names.index("a")
    a = values[A]
    va = valid[A]
    n = len(a)
    a1 = [a[i] * 100 for i in range(n) for j in range(n)]
    v1 = np.array([va[i] for i in range(n) for j in range(n)])
    a2 = [a[j] * 100 for i in range(n) for j in range(n)]
    v2 = np.array([va[j] for i in range(n) for j in range(n)])
    const = [999999999999999999] * (n * n)
    types2 = [T(DECIMAL, 8, 0, False, 18), T(DECIMAL, 8, 2, True, 18), T(DECIMAL, 8, 2, True, 18)]
    case = AggCase(types2, [], [(AGG_SUM, 0), (AGG_MIN, 1), (AGG_MAX, 2), (AGG_COUNT_STAR, -1)], global_agg=True)
    got = result_rows(backend.aggregate(case, [Batch(types2, [const, a1, a2], [None, v1, v2])]))
    return got, golden_rows(FIX["goldens"]["decimal_result/storage_sum"])


def _row_layout_cross():
    """t x t x t WHERE z.id < 4 (aggregate_row_layout), built on the host:
    columns x.id, y.id, x.n, z.a, x.v, y.v, z.v, z.f, y.f."""
    names, types, values, valid = table_columns(FIX["tables"]["aggregate_row_layout.t"])
    col = {n: (np.array(values[i]), valid[i] if valid[i] is not None else np.ones(len(values[i]), bool)) for i, n in enumerate(names)}
    ids = col["id"][0]
    zsel = np.nonzero(ids < 4)[0]
    n = len(ids)
    xi, yi, zi = np.meshgrid(np.arange(n), np.arange(n), zsel, indexing="ij")
    xi, yi, zi = xi.ravel(), yi.ravel(), zi.ravel()
    I64 = T(SIGNED, 8)
    I64N = T(SIGNED, 8, nullable=True)
    F64N = T(FLOAT, 8, nullable=True)
    data = [
        ("xid", I64, col["id"][0][xi], None), ("yid", I64, col["id"][0][yi], None),
        ("xn", I64N, col["n"][0][xi], col["n"][1][xi]), ("za", I64N, col["a"][0][zi], col["a"][1][zi]),
        ("xv", I64, col["v"][0][xi], None), ("yv", I64, col["v"][0][yi], None), ("zv", I64, col["v"][0][zi], None),
        ("zf", F64N, col["f"][0][zi].astype(np.float64), col["f"][1][zi]),
        ("yf", F64N, col["f"][0][yi].astype(np.float64), col["f"][1][yi]),
    ]
    return [d[0] for d in data], [d[1] for d in data], [d[2] for d in data], [d[3] for d in data]


def golden_estimated_groups(backend):
    """aggregate_row_layout/estimated_groups: GROUP BY x.id, y.id over
    t x t x t (z.id < 4) with COUNT(*), SUM(z.a), MIN(z.f), MAX(x.v + z.v),
    then a global COUNT/SUM/MIN/MAX over the 65,536 groups."""
    names, types, values, valid = _row_layout_cross()
    ix = {n: i for i, n in enumerate(names)}
    p = Prog(types)
    s = p.op(_abi().EXPR_ADD, p.col(ix["xv"]), p.col(ix["zv"]), T(SIGNED, 8))
    nin = len(types)
    inner = AggCase(types, [ix["xid"], ix["yid"]],
                    [(AGG_COUNT_STAR, -1), (AGG_SUM, ix["za"]), (AGG_MIN, ix["zf"]), (AGG_MAX, nin)],
                    computed=[(p, s, T(SIGNED, 8))])
    res = backend.aggregate(inner, [Batch(types, values, valid)], **({"capacity": 1024} if backend.name == "gpu" else {}))
    (i, _), (k, _), (c, _), (sa, sav), (nf, nfv), (xv, _) = res
    otypes = [T(SIGNED, 8), T(SIGNED, 8), T(SIGNED, 8), T(SIGNED, 8, nullable=True), T(FLOAT, 8, nullable=True), T(SIGNED, 8)]
    outer = AggCase(otypes, [], [(AGG_COUNT_STAR, -1), (AGG_SUM, 2), (AGG_SUM, 0), (AGG_SUM, 1), (AGG_SUM, 3),
                                 (AGG_MIN, 4), (AGG_MAX, 5)], global_agg=True)
    got = result_rows(backend.aggregate(outer, [Batch(otypes, [i, k, c, sa, nf, xv], [None, None, None, sav, nfv, None])]))
    return got, golden_rows(FIX["goldens"]["aggregate_row_layout/estimated_groups"])


def golden_checksum(backend):
    """aggregate_row_layout/checksum: GROUP BY d, y.id, x.n (d = x.id at
    scale 2; here DECIMAL(18,2), the golden casts to DECIMAL(20,2) whose
    values are the same) over t x t x t (z.id < 4) with COUNT(*), COUNT(z.a),
    SUM(z.a + y.v), MIN(z.a - x.v), MAX(z.a * y.v), SUM(x.v * y.v + z.v),
    MIN(z.f), MAX(y.f + z.f), SUM(z.f); then the global checksum over the
    65,536 groups."""
    names, types, values, valid = _row_layout_cross()
    ix = {n: i for i, n in enumerate(names)}
    d = values[ix["xid"]] * 100
    types = types + [T(DECIMAL, 8, 2, False, 18)]
    values = values + [d]
    valid = valid + [None]
    D = len(types) - 1
    I64N, F64N = T(SIGNED, 8, nullable=True), T(FLOAT, 8, nullable=True)
    E = _abi()

    def prog(fn):
        p = Prog(types)
        return (p, fn(p))

    sa = prog(lambda p: p.op(E.EXPR_ADD, p.col(ix["za"]), p.col(ix["yv"]), I64N))
    na = prog(lambda p: p.op(E.EXPR_SUB, p.col(ix["za"]), p.col(ix["xv"]), I64N))
    xa = prog(lambda p: p.op(E.EXPR_MUL, p.col(ix["za"]), p.col(ix["yv"]), I64N))
    sv = prog(lambda p: p.op(E.EXPR_ADD, p.op(E.EXPR_MUL, p.col(ix["xv"]), p.col(ix["yv"]), T(SIGNED, 8)), p.col(ix["zv"]), T(SIGNED, 8)))
    xf = prog(lambda p: p.op(E.EXPR_ADD, p.col(ix["yf"]), p.col(ix["zf"]), F64N, cls=E.REG_F64))
    nin = len(types)
    computed = [(sa[0], sa[1], I64N), (na[0], na[1], I64N), (xa[0], xa[1], I64N), (sv[0], sv[1], T(SIGNED, 8)),
                (xf[0], xf[1], F64N)]
    inner = AggCase(types, [D, ix["yid"], ix["xn"]],
                    [(AGG_COUNT_STAR, -1), (AGG_COUNT, ix["za"]), (AGG_SUM, nin), (AGG_MIN, nin + 1), (AGG_MAX, nin + 2),
                     (AGG_SUM, nin + 3), (AGG_MIN, ix["zf"]), (AGG_MAX, nin + 4), (AGG_SUM, ix["zf"])], computed=computed)
    res = backend.aggregate(inner, [Batch(types, values, valid)], **({"capacity": 1024} if backend.name == "gpu" else {}))
    cols = [c for c, _ in res]
    vals = [v for _, v in res]
    # d, k, n, c, ca, sa, na, xa, sv, nf, xf, sf
    otypes = [T(DECIMAL, 8, 2, False, 18), T(SIGNED, 8), I64N, T(SIGNED, 8), T(SIGNED, 8), I64N, I64N, I64N,
              T(SIGNED, 8), F64N, F64N, F64N]
    outer = AggCase(otypes, [], [(AGG_COUNT_STAR, -1), (AGG_SUM, 3), (AGG_SUM, 0), (AGG_SUM, 1), (AGG_SUM, 2), (AGG_COUNT, 2),
                                 (AGG_SUM, 4), (AGG_SUM, 5), (AGG_MIN, 6), (AGG_MAX, 7), (AGG_SUM, 8), (AGG_MIN, 9),
                                 (AGG_MAX, 10), (AGG_SUM, 11)], global_agg=True)
    got = result_rows(backend.aggregate(outer, [Batch(otypes, cols, vals)]))
    return got, golden_rows(FIX["goldens"]["aggregate_row_layout/checksum"])


def golden_aggregate_empty(backend, table, golden):
    """aggregate_empty/*_mixed: COUNT(*), COUNT(n), SUM(n), MIN(n), MAX(n),
    AVG(n) over a nullable i64 column, a global aggregate."""
    names, b = table_batch(table)
    case = AggCase(b.types, [], [(AGG_COUNT_STAR, -1), (AGG_COUNT, 0), (AGG_SUM, 0), (AGG_MIN, 0), (AGG_MAX, 0)], global_agg=True)
    r = result_rows(backend.aggregate(case, [b]))
    c, cn, s, mn, mx = r[0]
    (avg,) = backend_avg(backend, [s], [cn], 0, kind=SIGNED, width=8)
    got = [(c, cn, s, mn, mx, avg)]
    return got, golden_rows(FIX["goldens"][golden])


def golden_grouped_empty(backend):
    """aggregate_empty/grouped_empty: GROUP BY n over an empty table emits no rows."""
    names, b = table_batch("aggregate_empty.empty")
    case = AggCase(b.types, [0], [(AGG_COUNT_STAR, -1)])
    got = result_rows(backend.aggregate(case, [b]))
    return got, golden_rows(FIX["goldens"]["aggregate_empty/grouped_empty"])


def golden_outer_join(backend, golden, grp=None, right=False):
    """outer_join/*: c LEFT JOIN o ON c.ck = o.ok (o RIGHT JOIN c for
    right_join), optionally WHERE c.grp = grp (a probe-side selection);
    output (ck, cv, ok, ov), or (ok, ov, ck, cv) for the right join."""
    cn, cb = table_batch("outer_join.c")
    on, ob = table_batch("outer_join.o")
    if grp is not None:
        cb.selection = (np.array(cb.columns[cn.index("grp")]) == grp).astype(np.uint8)
    case = JoinCase(ob.types, cb.types, [on.index("ok")], [cn.index("ck")], join_type=JOIN_LEFT_OUTER,
                    probe_is_left=not right)
    pairs, rows = backend.join(case, [ob], cb)
    # project to the golden's columns
    if not right:
        sel = [0, 1, 3, 4]  # ck, cv, ok, ov  (probe c: ck cv grp | build o: ok ov flag)
    else:
        sel = [0, 1, 3, 4]  # ok, ov, ck, cv  (build o: ok ov flag | probe c: ck cv grp)
    got = sorted([tuple(r[i] for i in sel) for r in rows], key=sort_key)
    return got, sorted(golden_rows(FIX["goldens"][golden]), key=sort_key)


def golden_decimal_join(backend):
    """decimal_parquet/join: w LEFT JOIN m ON w.d2 = m.d2 (16-byte decimal
    keys, NULLs on both sides); the golden's m.name is checked through the
    matched build row."""
    wn, wb = table_batch("decimal_parquet.wide")
    mt = FIX["tables"]["decimal_parquet.wide_dim"]
    mn, mb = table_batch("decimal_parquet.wide_dim")
    names_m = [r[1] for r in mt["rows"]]
    case = JoinCase(mb.types, wb.types, [0], [wn.index("d2")], join_type=JOIN_LEFT_OUTER, probe_is_left=True)
    pairs, rows = backend.join(case, [mb], wb)
    got = []
    for a, bld in pairs:
        pr = int(a) + (1 << 32)
        d2 = wb.columns[wn.index("d2")][pr] if wb.valid[wn.index("d2")][pr] else None
        got.append((int(wb.columns[wn.index("id")][pr]), d2, names_m[bld] if bld >= 0 else None))
    got.sort(key=sort_key)
    return got, sorted(golden_rows(FIX["goldens"]["decimal_parquet/join"]), key=sort_key)


def golden_multiply_project(backend):
    """decimal_result/multiply (project): a2 * a2, a2 * b0, b0 * b0,
    a2 * 0.125, a2 * (1 - a2) per row; a2 = a as DECIMAL(10,2), b0 = b as
    DECIMAL(19,0)."""
    names, types, values, valid = _t_columns()
    A, B = names.index("a"), names.index("b")
    E = _abi()
    a2 = T(DECIMAL, 8, 2, True, 10)

    def build(fn, t):
        p = Prog(types)
        return (p, fn(p), t)

    cast_a = lambda p: p.scale_up(p.col(A), 2, a2)
    outs = [
        build(lambda p: p.op(E.EXPR_MUL, cast_a(p), cast_a(p), T(DECIMAL, 16, 4, True, 20)), T(DECIMAL, 16, 4, True, 20)),
        build(lambda p: p.op(E.EXPR_MUL, cast_a(p), p.col(B), T(DECIMAL, 16, 2, True, 29)), T(DECIMAL, 16, 2, True, 29)),
        build(lambda p: p.op(E.EXPR_MUL, p.col(B), p.col(B), T(DECIMAL, 16, 0, True, 38)), T(DECIMAL, 16, 0, True, 38)),
        build(lambda p: p.op(E.EXPR_MUL, cast_a(p), p.const(125, T(DECIMAL, 8, 3, False, 3)), T(DECIMAL, 8, 5, True, 13)), T(DECIMAL, 8, 5, True, 13)),
        build(lambda p: p.op(E.EXPR_MUL, cast_a(p), p.op(E.EXPR_SUB, p.const(100, T(DECIMAL, 8, 2, False, 3)), cast_a(p), T(DECIMAL, 8, 2, True, 11)), T(DECIMAL, 16, 4, True, 21)), T(DECIMAL, 16, 4, True, 21)),
    ]
    res = backend.project(types, outs, Batch(types, values, valid))
    ids = values[names.index("id")]
    got = []
    for r in range(len(ids)):
        row = [int(ids[r])]
        for vals, v in res:
            row.append(None if (v is not None and not v[r]) else int(vals[r]))
        got.append(tuple(row))
    return got, golden_rows(FIX["goldens"]["decimal_result/multiply"])


def golden_storage_error(backend, golden, table, column, scale, nullable):
    """decimal_result/storage_overflow (CAST(b AS DECIMAL(18,2)) over t) and
    storage_unsigned (CAST(u AS DECIMAL(18,0)) over u): a value with more
    than 18 digits cannot be stored as DECIMAL(18, s), an 8-byte decimal, so
    the query fails with the reference's runtime error."""
    names, types, values, valid = table_columns(FIX["tables"][table])
    c = names.index(column)
    out_t = T(DECIMAL, 8, scale, nullable, 18)
    p = Prog(types)
    r = p.scale_up(p.col(c), scale, out_t) if scale else p.col(c)
    try:
        backend.project(types, [(p, r, out_t)], Batch(types, values, valid))
        got = None
    except RuntimeError as e:
        got = "Runtime error: " + str(e)
    return got, FIX["errors"][golden]["text"]


def golden_storage_boundary(backend):
    """decimal_result/storage_boundary: CAST(a AS DECIMAL(9,2)) squared is a
    DECIMAL(18,4) held in 8 bytes and fits; the constant 999999999 *
    9999999999 is DECIMAL(19,0) in 16 bytes; CAST(a AS DECIMAL(10,2)) < b
    compares at scale 2. (The CASE column of the query is not an operator
    of this path and is not compared.)"""
    names, types, values, valid = _t_columns()
    A, B = names.index("a"), names.index("b")
    E = _abi()
    a2 = T(DECIMAL, 8, 2, True, 9)
    p18 = T(DECIMAL, 8, 4, True, 18)
    p19 = T(DECIMAL, 16, 0, False, 19)

    def build(fn, t):
        p = Prog(types)
        return (p, fn(p), t)

    outs = [
        build(lambda p: p.op(E.EXPR_MUL, p.scale_up(p.col(A), 2, a2), p.scale_up(p.col(A), 2, a2), p18), p18),
        build(lambda p: p.op(E.EXPR_MUL, p.const(999999999, T(DECIMAL, 8, 0, False, 9)),
                             p.const(9999999999, T(DECIMAL, 8, 0, False, 10)), p19), p19),
        build(lambda p: p.cmp(E.CMP_LT, p.scale_up(p.col(A), 2, T(DECIMAL, 8, 2, True, 10)), p.scale_up(p.col(B), 2, T(DECIMAL, 16, 2, True, 21))),
              T(3, 1, nullable=True)),
    ]
    res = backend.project(types, outs, Batch(types, values, valid))
    ids = values[names.index("id")]
    got = []
    for r in range(len(ids)):
        row = [int(ids[r])]
        for vals, v in res:
            row.append(None if (v is not None and not v[r]) else (bool(vals[r]) if isinstance(vals[r], (bool, np.bool_)) else int(vals[r])))
        got.append(tuple(row))
    want = [tuple(w[:4]) for w in golden_rows(FIX["goldens"]["decimal_result/storage_boundary"])]
    return got, want


AGG_GOLDENS = {
    "decimal_aggregates": golden_decimal_aggregates,
    "decimal_group_by": golden_decimal_group_by,
    "multiply_sum": golden_multiply_sum,
    "storage_sum": golden_storage_sum,
    "estimated_groups": golden_estimated_groups,
    "checksum": golden_checksum,
    "empty_mixed": lambda b: golden_aggregate_empty(b, "aggregate_empty.empty", "aggregate_empty/empty_mixed"),
    "nonempty_mixed": lambda b: golden_aggregate_empty(b, "aggregate_empty.values", "aggregate_empty/nonempty_mixed"),
    "all_null_mixed": lambda b: golden_aggregate_empty(b, "aggregate_empty.nulls", "aggregate_empty/all_null_mixed"),
    "grouped_empty": golden_grouped_empty,
}

JOIN_GOLDENS = {
    "duplicates": lambda b: golden_outer_join(b, "outer_join/duplicates"),
    "null_keys": lambda b: golden_outer_join(b, "outer_join/null_keys", grp=4),
    "all_matched": lambda b: golden_outer_join(b, "outer_join/all_matched", grp=1),
    "none_matched": lambda b: golden_outer_join(b, "outer_join/none_matched", grp=2),
    "right_join": lambda b: golden_outer_join(b, "outer_join/right_join", right=True),
    "decimal_join": golden_decimal_join,
}

PROJECT_GOLDENS = {
    "multiply": golden_multiply_project,
    "storage_boundary": golden_storage_boundary,
    "storage_overflow": lambda b: golden_storage_error(b, "decimal_result/storage_overflow", "decimal_result.t", "b", 2, True),
    "storage_unsigned": lambda b: golden_storage_error(b, "decimal_result/storage_unsigned", "decimal_result.u", "u", 0, False),
}


def random_agg_case(seed, n):
    rng = np.random.default_rng(seed)
    k1 = rng.integers(-3, 4, n).astype(np.int32)
    k1v = rng.random(n) > 0.1
    k2 = rng.integers(0, 50 + seed * 2000, n).astype(np.int64)
    vi = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    viv = rng.random(n) > 0.2
    dec = [int(x) * (10 ** 20) + int(y) for x, y in zip(rng.integers(-10 ** 9, 10 ** 9, n), rng.integers(0, 10 ** 9, n))]
    decv = rng.random(n) > 0.3
    vf = rng.normal(0, 1e3, n)
    d64 = rng.integers(-10 ** 12, 10 ** 12, n).astype(np.int64)
    types = [T(SIGNED, 4, nullable=True), T(SIGNED, 8), T(SIGNED, 8, nullable=True), T(DECIMAL, 16, 3, True, 38),
             T(FLOAT, 8), T(DECIMAL, 8, 2, False, 15)]
    case = AggCase(types, [0, 1], [(AGG_COUNT_STAR, -1), (AGG_COUNT, 2), (AGG_SUM, 2), (AGG_MIN, 2), (AGG_MAX, 2),
                                   (AGG_SUM, 3), (AGG_MIN, 3), (AGG_MAX, 3), (AGG_SUM, 4), (AGG_MIN, 4), (AGG_MAX, 4),
                                   (AGG_SUM, 5), (AGG_MIN, 5)])
    sel = (rng.random(n) > 0.15).astype(np.uint8)
    b = Batch(types, [k1, k2, vi, dec, vf, d64], [k1v, None, viv, decv, None, None], sel)
    return case, b
