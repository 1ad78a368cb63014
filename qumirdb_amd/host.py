"""Host-side mirror of the reference's operator interface for the path.

It names things as the reference does (TAggregateProcessor,
TRowJoinProcessor, agg_dispatch op-codes, TRowSet batches) and only calls the
HIP library through its C-ABI. Device memory comes from the library's pool;
numpy arrays are the host-side containers.

Decimal typing follows qdb/kernel/annotate_type.cpp: a sum or difference has
the larger scale and one more integer digit (:938-960), a product adds
precisions and scales, capped at 38 digits (DecimalProductSpec, :746-760);
storage is i64 up to precision 18, else i128 (qdb/plan/types/decimal.h:80-99).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import abi, check, lib

MAX_DECIMAL_PRECISION = 38


def init_device(device: int = 0) -> None:
    check(lib().qdb_gpu_init(device), "qdb_gpu_init")


@dataclass(frozen=True)
class ValueType:
    kind: int
    width: int
    scale: int = 0
    nullable: bool = False
    precision: int = 0  # decimals only

    def to_c(self) -> abi.QdbValueType:
        return abi.QdbValueType(self.kind, self.width, self.scale, self.nullable,
                                self.precision if self.kind == abi.DECIMAL else 0)

    def with_nullable(self, nullable: bool) -> "ValueType":
        return ValueType(self.kind, self.width, self.scale, nullable, self.precision)

    @property
    def np_dtype(self):
        if self.kind == abi.FLOAT:
            return np.float64 if self.width == 8 else np.float32
        if self.kind == abi.BOOL:
            return np.uint8
        if self.width == 16:
            return np.dtype([("lo", "<u8"), ("hi", "<i8")])
        signed = self.kind != abi.UNSIGNED
        return {1: np.int8 if signed else np.uint8, 2: np.int16 if signed else np.uint16,
                4: np.int32 if signed else np.uint32, 8: np.int64 if signed else np.uint64}[self.width]


def I8(nullable=False): return ValueType(abi.SIGNED, 1, nullable=nullable)
def I32(nullable=False): return ValueType(abi.SIGNED, 4, nullable=nullable)
def I64(nullable=False): return ValueType(abi.SIGNED, 8, nullable=nullable)
def F64(nullable=False): return ValueType(abi.FLOAT, 8, nullable=nullable)
def DATE32(nullable=False): return ValueType(abi.DATE, 4, nullable=nullable)


def DECIMAL(precision: int, scale: int, nullable: bool = False) -> ValueType:
    return ValueType(abi.DECIMAL, 8 if precision <= 18 else 16, scale, nullable, precision)


def integer_decimal(t: ValueType, literal: Optional[int] = None) -> ValueType:
    """IntegerDecimalSpec (annotate_type.cpp:818-845): a literal has its own
    digits, an integer column the digits of its width."""
    if literal is not None:
        return DECIMAL(max(1, len(str(abs(int(literal))))), 0)
    return DECIMAL({1: 3, 2: 5, 4: 10, 8: 19}.get(t.width, 38), 0)


def decimal_add_type(l: ValueType, r: ValueType) -> ValueType:
    scale = max(l.scale, r.scale)
    int_digits = max(l.precision - l.scale, r.precision - r.scale)
    return DECIMAL(min(MAX_DECIMAL_PRECISION, int_digits + scale + 1), scale, l.nullable or r.nullable)


def decimal_mul_type(l: ValueType, r: ValueType) -> ValueType:
    scale = l.scale + r.scale
    if scale > MAX_DECIMAL_PRECISION:
        raise ValueError(f"decimal product needs scale {scale}, above the DECIMAL maximum of 38")
    return DECIMAL(min(MAX_DECIMAL_PRECISION, l.precision + r.precision), scale, l.nullable or r.nullable)


def decimal_sum_type(t: ValueType) -> ValueType:
    """SUM of a DECIMAL keeps the scale at precision 38 (aggregate.cpp:40-44)."""
    return DECIMAL(MAX_DECIMAL_PRECISION, t.scale, t.nullable)


class Program:
    """Builder of an expression program (QdbExprProgram) with decimal typing.

    Methods return value handles; every operand is consumed by the operation
    that uses it, so its register is reused (8 registers hold any tree of up
    to 8 live values)."""

    def __init__(self, schema: Sequence[ValueType]):
        self.schema = list(schema)
        self.insns: list[abi.QdbExprInsn] = []
        self.types: list[ValueType] = []   # per value
        self.regs: list[int] = []          # per value
        self.free = list(range(abi.MAX_REGS))

    def _emit(self, op, cls, typ, a=None, b=None, column=0, imm_lo=0, imm_hi=0) -> int:
        ra = self.regs[a] if a is not None else 0
        rb = self.regs[b] if b is not None else 0
        for v in (a, b):
            if v is not None and self.regs[v] not in self.free:
                self.free.append(self.regs[v])
        if not self.free:
            raise ValueError("expression needs more than 8 live registers")
        self.free.sort()
        dst = self.free.pop(0)
        ins = abi.QdbExprInsn()
        ins.Op, ins.Dst, ins.A, ins.B, ins.Class = op, dst, ra, rb, cls
        ins.Column, ins.ImmLo, ins.ImmHi = column, imm_lo, imm_hi
        self.insns.append(ins)
        self.types.append(typ)
        self.regs.append(dst)
        return len(self.types) - 1

    def col(self, index: int) -> int:
        t = self.schema[index]
        cls = abi.REG_F64 if t.kind == abi.FLOAT else abi.REG_BOOL if t.kind == abi.BOOL else abi.REG_INT
        return self._emit(abi.EXPR_COLUMN, cls, t, column=index)

    def const_int(self, value: int, typ: Optional[ValueType] = None) -> int:
        v = int(value) & ((1 << 128) - 1)
        lo, hi = v & ((1 << 64) - 1), v >> 64
        to_s64 = lambda x: x - (1 << 64) if x >= (1 << 63) else x
        return self._emit(abi.EXPR_CONST_INT, abi.REG_INT, typ or integer_decimal(I64(), value),
                          imm_lo=to_s64(lo), imm_hi=to_s64(hi))

    def const_decimal(self, unscaled: int, precision: int, scale: int) -> int:
        return self.const_int(unscaled, DECIMAL(precision, scale))

    def const_f64(self, value: float) -> int:
        bits = int(np.array([value], dtype=np.float64).view(np.int64)[0])
        return self._emit(abi.EXPR_CONST_F64, abi.REG_F64, F64(), imm_lo=bits)

    def cast_decimal(self, v: int, precision: int, scale: int) -> int:
        """CAST(integer or decimal AS DECIMAL(p, s)) for s >= the input scale."""
        t = self.types[v]
        cur = t.scale if t.kind == abi.DECIMAL else 0
        if scale < cur:
            raise ValueError("narrowing decimal casts are not supported here")
        return self._emit(abi.EXPR_SCALE_UP, abi.REG_INT, DECIMAL(precision, scale, t.nullable), a=v, imm_lo=scale - cur)

    def _align(self, v: int, scale: int) -> int:
        t = self.types[v]
        if t.kind == abi.DECIMAL and t.scale < scale:
            nt = DECIMAL(min(38, t.precision + scale - t.scale), scale, t.nullable)
            return self._emit(abi.EXPR_SCALE_UP, abi.REG_INT, nt, a=v, imm_lo=scale - t.scale)
        return v

    def _as_decimal(self, v: int) -> ValueType:
        t = self.types[v]
        if t.kind == abi.DECIMAL:
            return t
        return integer_decimal(t)

    def _arith(self, op, a: int, b: int) -> int:
        ta, tb = self.types[a], self.types[b]
        if ta.kind == abi.FLOAT or tb.kind == abi.FLOAT:
            return self._emit(op, abi.REG_F64, F64(nullable=ta.nullable or tb.nullable), a=a, b=b)
        if ta.kind != abi.DECIMAL and tb.kind != abi.DECIMAL:
            return self._emit(op, abi.REG_INT, I64(nullable=ta.nullable or tb.nullable), a=a, b=b)
        da, db = self._as_decimal(a), self._as_decimal(b)
        if op == abi.EXPR_MUL:
            return self._emit(op, abi.REG_INT, decimal_mul_type(da, db), a=a, b=b)
        rt = decimal_add_type(da, db)
        a2 = self._align(a, rt.scale)
        b2 = self._align(b, rt.scale)
        return self._emit(op, abi.REG_INT, rt, a=a2, b=b2)

    def add(self, a, b): return self._arith(abi.EXPR_ADD, a, b)
    def sub(self, a, b): return self._arith(abi.EXPR_SUB, a, b)
    def mul(self, a, b): return self._arith(abi.EXPR_MUL, a, b)

    def cmp(self, op: int, a: int, b: int) -> int:
        ta, tb = self.types[a], self.types[b]
        if ta.kind == abi.DECIMAL or tb.kind == abi.DECIMAL:
            scale = max(self._as_decimal(a).scale, self._as_decimal(b).scale)
            a = self._align(a, scale)
            b = self._align(b, scale)
        nullable = ta.nullable or tb.nullable
        return self._emit(abi.EXPR_CMP, abi.REG_BOOL, ValueType(abi.BOOL, 1, nullable=nullable), a=a, b=b, imm_lo=op)

    def and_(self, a, b): return self._emit(abi.EXPR_AND, abi.REG_BOOL, ValueType(abi.BOOL, 1, nullable=True), a=a, b=b)
    def or_(self, a, b): return self._emit(abi.EXPR_OR, abi.REG_BOOL, ValueType(abi.BOOL, 1, nullable=True), a=a, b=b)
    def not_(self, a): return self._emit(abi.EXPR_NOT, abi.REG_BOOL, ValueType(abi.BOOL, 1, nullable=True), a=a)
    def is_null(self, a): return self._emit(abi.EXPR_IS_NULL, abi.REG_BOOL, ValueType(abi.BOOL, 1), a=a)

    def to_f64(self, a) -> int:
        """CAST(decimal or integer AS DOUBLE): the nearest double, ties to
        even (qdb_decimal_to_f64, qdb/modules/qumirdb.oz:983-1046)."""
        t = self.types[a]
        return self._emit(abi.EXPR_TO_F64, abi.REG_F64, ValueType(abi.FLOAT, 8, nullable=t.nullable), a=a,
                          imm_lo=t.scale if t.kind == abi.DECIMAL else 0)

    def div(self, a, b) -> int:
        """DOUBLE division (IEEE)."""
        nullable = self.types[a].nullable or self.types[b].nullable
        return self._emit(abi.EXPR_DIV, abi.REG_F64, ValueType(abi.FLOAT, 8, nullable=nullable), a=a, b=b)

    def avg(self, sum_col: int, count_col: int) -> int:
        """AVG as the reference lowers it (qdb/plan/sql_lowering.cpp:799-814):
        SUM / COUNT with both converted to DOUBLE. A group whose COUNT is 0
        has a NULL SUM, so its AVG is NULL."""
        return self.div(self.to_f64(self.col(sum_col)), self.to_f64(self.col(count_col)))

    def build(self, result: int) -> abi.QdbExprProgram:
        p = abi.QdbExprProgram()
        if len(self.insns) > abi.MAX_INSNS:
            raise ValueError("expression has too many instructions")
        p.NumInsns = len(self.insns)
        p.Result = self.regs[result]
        for k, ins in enumerate(self.insns):
            p.Insns[k] = ins
        return p

    def result_type(self, result: int) -> ValueType:
        return self.types[result]


def schema_c(types: Sequence[ValueType]) -> abi.QdbColumnSchema:
    s = abi.QdbColumnSchema()
    s.NumColumns = len(types)
    for i, t in enumerate(types):
        s.Types[i] = t.to_c()
    return s


class DeviceBuffer:
    """A block of device memory from the library's pool."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        self.ptr = lib().qdb_gpu_alloc(max(self.nbytes, 1))
        if not self.ptr:
            raise MemoryError(f"qdb_gpu_alloc({nbytes}) failed")

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(lib().qdb_gpu_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes), "memcpy h2d")
        return b

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            check(lib().qdb_gpu_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "memcpy d2h")
        return out

    def zero(self) -> None:
        check(lib().qdb_gpu_memset(self.ptr, 0, self.nbytes), "memset")

    def free(self) -> None:
        if self.ptr:
            lib().qdb_gpu_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def pack_bitmap(valid: np.ndarray) -> np.ndarray:
    """Arrow LSB-first validity bitmap of a bool array."""
    return np.packbits(np.asarray(valid, dtype=np.uint8), bitorder="little")


def unpack_bitmap(bits: np.ndarray, n: int) -> np.ndarray:
    return np.unpackbits(bits, bitorder="little", count=n).astype(bool)


class DeviceRowSet:
    """A TRowSet whose column buffers live in HBM.

    columns: numpy arrays (bool columns as 0/1 arrays, packed here; 16-byte
    decimals as the structured (lo, hi) dtype or as Python ints), or
    DeviceBuffer objects; valid: per column None or a bool array.
    """

    def __init__(self, types: Sequence[ValueType], columns, valid=None, selection=None, rows: Optional[int] = None):
        self.types = list(types)
        self.buffers: list = []
        n = rows
        cols = (abi.TColumn * max(1, len(self.types)))()
        for i, (t, data) in enumerate(zip(self.types, columns)):
            if isinstance(data, DeviceBuffer):
                buf = data
            else:
                arr = data
                if t.kind == abi.BOOL:
                    arr = pack_bitmap(np.asarray(data).astype(bool))
                    n = len(data) if n is None else n
                elif t.width == 16 and not (isinstance(data, np.ndarray) and data.dtype.names):
                    arr = i128_array(data)
                    n = len(data) if n is None else n
                else:
                    arr = np.ascontiguousarray(np.asarray(data, dtype=t.np_dtype))
                    n = len(arr) if n is None else n
                buf = DeviceBuffer.from_numpy(arr)
            self.buffers.append(buf)
            cols[i].Data = buf.ptr
            v = None if valid is None else valid[i]
            if v is not None:
                mb = DeviceBuffer.from_numpy(pack_bitmap(v))
                self.buffers.append(mb)
                cols[i].Mask = mb.ptr
        self.rows = int(n or 0)
        self.cols = cols
        self.rs = abi.TRowSet()
        self.rs.Columns = C.cast(cols, C.POINTER(abi.TColumn))
        self.rs.ColumnCount = len(self.types)
        self.rs.RowCount = self.rows
        self.rs.RefCount = 1
        self.sel_buf = None
        if selection is not None:
            self.sel_buf = DeviceBuffer.from_numpy(np.asarray(selection, dtype=np.uint8))
            self.rs.Selection = self.sel_buf.ptr

    def view(self, r0: int, r1: int) -> "DeviceRowSet":
        """Rows [r0, r1) of this batch without a copy: data pointers
        advanced, bit-packed data and validity by their bit offsets."""
        if not 0 <= r0 <= r1 <= self.rows:
            raise ValueError(f"rows [{r0}, {r1}) outside [0, {self.rows})")
        v = DeviceRowSet.__new__(DeviceRowSet)
        v.types = self.types
        v.buffers = [self]  # keeps the parent's buffers alive
        cols = (abi.TColumn * max(1, len(self.types)))()
        for i, t in enumerate(self.types):
            c = self.cols[i]
            if t.kind == abi.BOOL:
                cols[i].Data = c.Data
                cols[i].DataBitOffset = c.DataBitOffset + r0
            else:
                cols[i].Data = (c.Data or 0) + r0 * t.width
            if c.Mask:
                cols[i].Mask = c.Mask
                cols[i].MaskBitOffset = c.MaskBitOffset + r0
        v.rows = r1 - r0
        v.cols = cols
        v.rs = abi.TRowSet()
        v.rs.Columns = C.cast(cols, C.POINTER(abi.TColumn))
        v.rs.ColumnCount = len(self.types)
        v.rs.RowCount = v.rows
        v.rs.RefCount = 1
        v.sel_buf = None
        if self.rs.Selection:
            v.rs.Selection = self.rs.Selection + r0
        return v

    def set_selection_buffer(self, buf: Optional[DeviceBuffer]) -> None:
        self.sel_buf = buf
        self.rs.Selection = buf.ptr if buf is not None else None

    @property
    def ptr(self):
        return C.byref(self.rs)


class RowSetView:
    """A device TRowSet made by the library (an aggregate's or a join's
    output), seen as a kernel input; the owner releases it."""

    def __init__(self, rs: abi.TRowSet, types: Sequence[ValueType]):
        self.rs = rs
        self.types = list(types)
        self.rows = int(rs.RowCount)

    @property
    def ptr(self):
        return C.byref(self.rs)


def i128_array(values) -> np.ndarray:
    dt = np.dtype([("lo", "<u8"), ("hi", "<i8")])
    out = np.empty(len(values), dtype=dt)
    for i, v in enumerate(values):
        v = int(v) if v is not None else 0
        u = v & ((1 << 128) - 1)
        out[i] = (u & ((1 << 64) - 1), (u >> 64) - (1 << 64) if (u >> 127) else (u >> 64))
    return out


def i128_values(arr: np.ndarray) -> list[int]:
    return [int(h) * (1 << 64) + int(l) for l, h in zip(arr["lo"].tolist(), arr["hi"].tolist())]


def download_rowset(rs: abi.TRowSet, types: Sequence[ValueType]):
    """Copies a device TRowSet to host: a list of (values, valid) per column;
    16-byte columns come back as Python ints, bools as bool arrays."""
    n = rs.RowCount
    tarr = (abi.QdbValueType * max(1, len(types)))(*[t.to_c() for t in types])
    host = abi.TRowSet()
    check(lib().qdb_gpu_download_rowset(C.byref(rs), tarr, C.byref(host)), "download_rowset")
    out = []
    try:
        for i, t in enumerate(types):
            col = host.Columns[i]
            if t.kind == abi.BOOL:
                raw = np.ctypeslib.as_array(C.cast(col.Data, C.POINTER(C.c_uint8)), shape=((n + 7) // 8,)).copy() if n else np.zeros(0, np.uint8)
                vals = unpack_bitmap(raw, n)
            elif n == 0:
                vals = np.zeros(0, dtype=t.np_dtype)
            else:
                nbytes = n * t.width
                raw = np.ctypeslib.as_array(C.cast(col.Data, C.POINTER(C.c_uint8)), shape=(nbytes,)).copy()
                vals = raw.view(t.np_dtype)
            if t.width == 16 and t.kind != abi.BOOL:
                vals = i128_values(vals)
            valid = None
            if col.Mask:
                mraw = np.ctypeslib.as_array(C.cast(col.Mask, C.POINTER(C.c_uint8)), shape=((n + 7) // 8,)).copy() if n else np.zeros(0, np.uint8)
                valid = unpack_bitmap(mraw, n)
            out.append((vals, valid))
    finally:
        lib().qdb_gpu_rowset_release(C.byref(host))
    return out


class FilterKernel:
    def __init__(self, types: Sequence[ValueType], program: Program, result: int):
        spec = abi.QdbFilterSpec()
        spec.Input = schema_c(types)
        spec.Predicate = program.build(result)
        fn = C.c_void_p()
        self.slot = check(lib().qdb_gpu_compile_filter(C.byref(spec), C.byref(fn)), "compile_filter")
        self.fn = abi.FILTER_FN(fn.value)

    def __call__(self, rowset: DeviceRowSet, input_sel: Optional[DeviceBuffer] = None) -> DeviceBuffer:
        out = DeviceBuffer(max(rowset.rows, 1))
        view = abi.TRowSet.from_buffer_copy(rowset.rs)
        view.Selection = out.ptr
        self.fn(C.byref(view), None, None, input_sel.ptr if input_sel is not None else None)
        err = lib().qdb_gpu_last_error()
        if err:
            raise RuntimeError(f"filter: {err.decode()}")
        return out


class ProjectKernel:
    def __init__(self, types: Sequence[ValueType], outputs: Sequence[tuple[Program, int]]):
        spec = abi.QdbProjectSpec()
        spec.Input = schema_c(types)
        spec.NumOutputs = len(outputs)
        self.out_types = []
        for k, (prog, res) in enumerate(outputs):
            t = prog.result_type(res)
            self.out_types.append(t)
            spec.OutputTypes[k] = t.to_c()
            spec.Outputs[k] = prog.build(res)
        fn = C.c_void_p()
        self.slot = check(lib().qdb_gpu_compile_project(C.byref(spec), C.byref(fn)), "compile_project")
        self.fn = abi.PROJECT_FN(fn.value)

    def __call__(self, rowset: DeviceRowSet, input_sel: Optional[DeviceBuffer] = None):
        n = max(rowset.rows, 1)
        k = len(self.out_types)
        bufs = []
        ptrs = (C.c_void_p * (2 * k))()
        for i, t in enumerate(self.out_types):
            b = DeviceBuffer((n + 7) // 8 + 8 if t.kind == abi.BOOL else n * t.width)
            b.zero()
            bufs.append(b)
            ptrs[i] = b.ptr
            if t.nullable:
                m = DeviceBuffer((n + 7) // 8 + 8)
                m.zero()
                bufs.append(m)
                ptrs[k + i] = m.ptr
        sel = input_sel.ptr if input_sel is not None else None
        self.fn(C.byref(rowset.rs), ptrs, None, None, sel)
        if lib().qdb_gpu_take_row_error():
            # A row failed (qdb_decimal64_overflow): as the reference's
            # project operator does (qdb/exec/project_exec.cpp:71-81), the
            # rows of the selection are evaluated again and only a failure
            # among them is an error. Here the first pass already covered
            # the selection when one was given.
            if sel is None and rowset.rs.Selection:
                self.fn(C.byref(rowset.rs), ptrs, None, None, rowset.rs.Selection)
                if not lib().qdb_gpu_take_row_error():
                    return bufs, ptrs
            raise RuntimeError(f"project: {lib().qdb_gpu_last_error().decode()}")
        err = lib().qdb_gpu_last_error()
        if err:
            raise RuntimeError(f"project: {err.decode()}")
        return bufs, ptrs


@dataclass
class AggSpec:
    """A group-by: key columns, (func, column) aggregates, optional fused
    predicate and computed argument columns."""
    types: Sequence[ValueType]
    keys: Sequence[int]
    aggs: Sequence[tuple[int, int]]
    empty_input_groups: Sequence[int] = ()
    output_on_host: bool = False
    predicate: Optional[tuple[Program, int]] = None
    computed: Sequence[tuple[Program, int]] = field(default_factory=list)

    def to_c(self) -> abi.QdbAggregateSpec:
        s = abi.QdbAggregateSpec()
        s.Input = schema_c(self.types)
        s.NumKeys = len(self.keys)
        for i, k in enumerate(self.keys):
            s.KeyColumns[i] = k
        s.NumAggs = len(self.aggs)
        for i, (f, c) in enumerate(self.aggs):
            s.Aggs[i] = abi.QdbAggregateFn(f, c)
        s.NumEmptyInputGroups = len(self.empty_input_groups)
        for i, g in enumerate(self.empty_input_groups):
            s.EmptyInputGroups[i] = g
        s.OutputOnHost = int(self.output_on_host)
        if self.predicate is not None:
            s.HasPredicate = 1
            s.Predicate = self.predicate[0].build(self.predicate[1])
        s.NumComputed = len(self.computed)
        for k, (prog, res) in enumerate(self.computed):
            s.ComputedTypes[k] = prog.result_type(res).to_c()
            s.Computed[k] = prog.build(res)
        return s

    def all_types(self) -> list[ValueType]:
        return list(self.types) + [p.result_type(r) for p, r in self.computed]


def combine_spec(spec: AggSpec, partial_types: Sequence[ValueType]) -> AggSpec:
    """The final aggregate over a partial aggregate's outputs
    (BuildAggregateCombineOperator, qdb/exec/planner_helpers.cpp:191-209):
    the same keys, COUNT and COUNT(*) become SUM, SUM/MIN/MAX stay, each
    reading its own partial output column."""
    nk = len(spec.keys)
    aggs = [(abi.AGG_SUM if f in (abi.AGG_COUNT, abi.AGG_COUNT_STAR) else f, nk + i) for i, (f, _) in enumerate(spec.aggs)]
    return AggSpec(types=list(partial_types), keys=list(range(nk)), aggs=aggs)


class AggregateKernels:
    """Compiled agg_dispatch / agg_finish_rowset of one spec."""

    def __init__(self, spec: AggSpec):
        self.spec = spec
        self.c_spec = spec.to_c()
        self.k = abi.QdbAggregateKernels()
        check(lib().qdb_gpu_compile_aggregate(C.byref(self.c_spec), C.byref(self.k)), "compile_aggregate")
        self.dispatch = abi.AGG_DISPATCH_FN(self.k.Dispatch)
        self.finish = abi.AGG_FINISH_FN(self.k.FinishRowSet)
        self.out_types = [ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in
                          self.k.OutputTypes[: self.k.NumOutputColumns]]

    def release(self):
        lib().qdb_gpu_release_slot(abi.SLOT_AGGREGATE, self.k.Slot)


class TAggregateProcessor:
    """qdb/exec/aggregate_exec.cpp:52-156 over the C-ABI."""

    def __init__(self, kernels: AggregateKernels, initial_capacity: int = 8, row_layout: bool = False,
                 initial_groups: int = 0):
        self.kernels = kernels
        self.p = lib().qdb_gpu_agg_processor_new(C.byref(kernels.k), initial_capacity, int(row_layout), initial_groups)
        if not self.p:
            raise RuntimeError(f"aggregate processor: {(lib().qdb_gpu_exec_last_error() or b'').decode()}")

    def add(self, batch: DeviceRowSet) -> None:
        if lib().qdb_gpu_agg_processor_add(self.p, batch.ptr) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    @staticmethod
    def add_batched(processors: Sequence["TAggregateProcessor"], batches: Sequence[DeviceRowSet]) -> None:
        """processors[i].add(batches[i]) for every i, as ONE batched launch
        (all processors of one compiled aggregate)."""
        n = len(processors)
        if n != len(batches):
            raise ValueError("one batch per processor")
        procs = (C.c_void_p * max(n, 1))(*[p.p for p in processors])
        rows = (C.POINTER(abi.TRowSet) * max(n, 1))(*[C.pointer(b.rs) for b in batches])
        if lib().qdb_gpu_agg_processor_add_batched(procs, rows, n) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    @staticmethod
    def _finish_batched_raw(processors: Sequence["TAggregateProcessor"]):
        n = len(processors)
        procs = (C.c_void_p * max(n, 1))(*[p.p for p in processors])
        outs = (abi.TRowSet * max(n, 1))()
        if lib().qdb_gpu_agg_processor_finish_batched(procs, outs, n) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        return outs

    @staticmethod
    def finish_batched(processors: Sequence["TAggregateProcessor"]):
        """processors[i].finish() for every i, with one synchronization for
        all of them (all processors of one compiled aggregate)."""
        outs = TAggregateProcessor._finish_batched_raw(processors)
        res = []
        for i, p in enumerate(processors):
            try:
                res.append(download_rowset(outs[i], p.kernels.out_types))
            finally:
                lib().qdb_gpu_rowset_release(C.byref(outs[i]))
        return res

    @staticmethod
    def finish_in_hbm_batched(processors: Sequence["TAggregateProcessor"]) -> List[int]:
        """finish_in_hbm() of every processor with one synchronization;
        returns the group counts."""
        outs = TAggregateProcessor._finish_batched_raw(processors)
        counts = []
        for i in range(len(processors)):
            counts.append(int(outs[i].RowCount))
            lib().qdb_gpu_rowset_release(C.byref(outs[i]))
        return counts

    def groups(self) -> int:
        return lib().qdb_gpu_agg_processor_groups(self.p)

    def finish(self):
        out = abi.TRowSet()
        if lib().qdb_gpu_agg_processor_finish(self.p, C.byref(out)) < 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        try:
            return download_rowset(out, self.kernels.out_types)
        finally:
            lib().qdb_gpu_rowset_release(C.byref(out))

    def finish_raw(self) -> abi.TRowSet:
        """Finishes into a device-resident rowset (adopted: release it with
        lib().qdb_gpu_rowset_release)."""
        out = abi.TRowSet()
        if lib().qdb_gpu_agg_processor_finish(self.p, C.byref(out)) < 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        return out

    def finish_in_hbm(self) -> int:
        """Finishes into a device-resident rowset, as a downstream GPU
        operator would take it, releases it and returns the group count."""
        out = abi.TRowSet()
        if lib().qdb_gpu_agg_processor_finish(self.p, C.byref(out)) < 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        n = int(out.RowCount)
        lib().qdb_gpu_rowset_release(C.byref(out))
        return n

    def close(self):
        if self.p:
            lib().qdb_gpu_agg_processor_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class JoinSpec:
    build_types: Sequence[ValueType]
    probe_types: Sequence[ValueType]
    build_keys: Sequence[int]
    probe_keys: Sequence[int]
    join_type: int = abi.JOIN_INNER
    probe_is_left: bool = True
    fused: Sequence[tuple[int, int]] = field(default_factory=list)
    table_mode: int = abi.JOIN_TABLE_AUTO

    def to_c(self) -> abi.QdbRowJoinSpec:
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


class RowJoinKernels:
    def __init__(self, spec: JoinSpec):
        self.spec = spec
        self.c_spec = spec.to_c()
        self.k = abi.QdbRowJoinKernels()
        check(lib().qdb_gpu_compile_row_join(C.byref(self.c_spec), C.byref(self.k)), "compile_row_join")
        self.dispatch = abi.JOIN_DISPATCH_FN(self.k.Dispatch)
        self.materialize = abi.JOIN_MATERIALIZE_FN(self.k.Materialize)
        self.out_types = [ValueType(t.Kind, t.Width, t.Scale, bool(t.Nullable)) for t in
                          self.k.OutputTypes[: self.k.NumOutputColumns]]

    def release(self):
        lib().qdb_gpu_release_slot(abi.SLOT_JOIN, self.k.Slot)


class TRowJoinProcessor:
    """qdb/exec/row_join_exec.cpp:22-276 over the C-ABI (build side fixed)."""

    def __init__(self, kernels: RowJoinKernels, build_is_left: bool = False, output_batch_rows: int = 16384):
        self.kernels = kernels
        self.p = lib().qdb_gpu_join_processor_new(C.byref(kernels.k), int(build_is_left), output_batch_rows)
        if not self.p:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    def _ok(self, rc):
        if rc < 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        return rc

    def append(self, batch: DeviceRowSet):
        self._ok(lib().qdb_gpu_join_processor_append(self.p, batch.ptr))

    def build(self):
        self._ok(lib().qdb_gpu_join_processor_build(self.p))

    def probe(self, batch: DeviceRowSet):
        """Probes and materializes; returns the output batches (host copies)."""
        self._ok(lib().qdb_gpu_join_processor_probe(self.p, batch.ptr))
        outs = []
        while True:
            out = abi.TRowSet()
            if not lib().qdb_gpu_join_processor_next(self.p, C.byref(out)):
                break
            try:
                outs.append(download_rowset(out, self.kernels.out_types))
            finally:
                lib().qdb_gpu_rowset_release(C.byref(out))
        return outs

    def probe_pairs(self, batch: DeviceRowSet) -> np.ndarray:
        """Probes only; returns the pairs as an (n, 2) int64 array."""
        n = self._ok(lib().qdb_gpu_join_processor_probe_only(self.p, batch.ptr))
        pb = lib().qdb_gpu_join_processor_pairs(self.p).contents
        out = np.empty(2 * n, dtype=np.int64)
        if n:
            check(lib().qdb_gpu_memcpy_d2h(out.ctypes.data, pb.Data, out.nbytes), "pairs d2h")
        return out.reshape(n, 2)

    @staticmethod
    def build_batched(procs: Sequence["TRowJoinProcessor"]) -> None:
        """build() of every processor (one compiled join) with one
        synchronization for all their dense tables."""
        arr = (C.c_void_p * max(1, len(procs)))(*[p.p for p in procs])
        if lib().qdb_gpu_join_processor_build_batched(arr, len(procs)) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    @staticmethod
    def probe_batched(procs: Sequence["TRowJoinProcessor"], batches: Sequence[DeviceRowSet]) -> None:
        """Probe-only of batches[i] by procs[i] (one compiled join) in one
        persistent launch; each processor's pairs stay readable
        (pairs, pairs_checksum, fused_results)."""
        n = len(procs)
        arr = (C.c_void_p * max(1, n))(*[p.p for p in procs])
        if n != len(batches):
            raise ValueError("one batch per processor")
        bs = (C.POINTER(abi.TRowSet) * max(1, n))(*[C.pointer(b.rs) for b in batches])
        if lib().qdb_gpu_join_processor_probe_batched(arr, bs, n) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    def probe_only(self, batch: DeviceRowSet) -> int:
        """Probes without materializing; returns the pair count (the pairs
        stay on the device until the next probe)."""
        return self._ok(lib().qdb_gpu_join_processor_probe_only(self.p, batch.ptr))

    def pairs_checksum(self) -> tuple:
        """(sum, xor) of a mix of every pair of the last probe-only call,
        computed on the device (equal for equal pair sets in any order)."""
        pb = lib().qdb_gpu_join_processor_pairs(self.p).contents
        out = (C.c_uint64 * 2)()
        check(lib().qdb_gpu_pairs_checksum(pb.Data, int(pb.Count), out), "pairs_checksum")
        return int(out[0]), int(out[1])

    def probe_detached(self, batches: Sequence[DeviceRowSet], pair_buffers: Optional[Sequence] = None):
        """Probes every batch against this processor's built table in one
        replay of the calling thread's captured pipeline, without touching
        the processor's own pairs or fused totals. pair_buffers: one
        abi.QdbPairBuffer per batch (or None: pairs counted, not written).
        Returns [(pair count, fused results)] per batch."""
        n = len(batches)
        bs = (C.POINTER(abi.TRowSet) * max(1, n))(*[C.pointer(b.rs) for b in batches])
        pbs = None
        if pair_buffers is not None:
            pbs = (C.POINTER(abi.QdbPairBuffer) * max(1, n))(*[C.pointer(p) if p is not None else C.POINTER(abi.QdbPairBuffer)()
                                                               for p in pair_buffers])
        counts = (C.c_int64 * max(1, n))()
        nf = len(self.kernels.spec.fused)
        fused = (abi.QdbFusedResult * max(1, n * nf))()
        arena = lib().qdb_gpu_join_processor_arena(self.p)
        check(lib().qdb_gpu_join_probe_detached(C.byref(self.kernels.k), arena, bs, pbs, n, counts, fused), "join_probe_detached")
        out = []
        for i in range(n):
            res = []
            for r in fused[i * nf:(i + 1) * nf]:
                v = (int(r.Hi) << 64) + (int(r.Lo) & ((1 << 64) - 1))
                res.append({"value": v, "count": int(r.Count), "valid": bool(r.Valid)})
            out.append((int(counts[i]), res))
        return out

    def table_info(self) -> dict:
        """The built table: dense or hash, its key range, packed ordinal bits."""
        v = (C.c_int64 * 6)()
        lib().qdb_gpu_join_table_info(lib().qdb_gpu_join_processor_arena(self.p), v, 6)
        return {"built": bool(v[0]), "dense": bool(v[1]), "range": int(v[2]), "packed_bits": int(v[3]), "rows": int(v[4]),
                "appended": int(v[5])}

    def pair_count(self) -> int:
        return int(lib().qdb_gpu_join_processor_pairs(self.p).contents.Count)

    def pairs(self) -> np.ndarray:
        """The pairs of the last probe-only call as an (n, 2) int64 array."""
        pb = lib().qdb_gpu_join_processor_pairs(self.p).contents
        n = int(pb.Count)
        out = np.empty(2 * n, dtype=np.int64)
        if n:
            check(lib().qdb_gpu_memcpy_d2h(out.ctypes.data, pb.Data, out.nbytes), "pairs d2h")
        return out.reshape(n, 2)

    def fused_results(self):
        arena = lib().qdb_gpu_join_processor_arena(self.p)
        n = len(self.kernels.spec.fused)
        res = (abi.QdbFusedResult * max(1, n))()
        lib().qdb_gpu_join_fused_results(arena, res, n)
        out = []
        for r in res[:n]:
            v = (int(r.Hi) << 64) + (int(r.Lo) & ((1 << 64) - 1))
            out.append({"value": v, "count": int(r.Count), "valid": bool(r.Valid)})
        return out

    def close(self):
        if self.p:
            lib().qdb_gpu_join_processor_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RuntimeFilter:
    """The runtime join filter over the C-ABI (qdb_gpu_rf_*): built from the
    selected build rows' join-key hashes, published once, probed per probe
    batch (TRuntimeFilter, qdb/exec/runtime_filter.cpp)."""

    def __init__(self, max_keys: int = 1 << 14, bloom_bytes: int = 2 << 20):
        self.f = lib().qdb_gpu_rf_new(max_keys, bloom_bytes)
        if not self.f:
            raise RuntimeError(f"runtime filter: {lib().qdb_gpu_last_error().decode()}")

    def build(self, hashes: DeviceBuffer, count: int, selection: Optional[DeviceBuffer] = None) -> None:
        check(lib().qdb_gpu_rf_build(self.f, hashes.ptr, selection.ptr if selection is not None else None, count), "rf_build")

    def publish(self) -> int:
        return check(lib().qdb_gpu_rf_publish(self.f), "rf_publish")

    def probe(self, hashes: DeviceBuffer, count: int, selection: Optional[DeviceBuffer] = None):
        """(selection buffer of the kept rows, kept count)."""
        out = DeviceBuffer(max(count, 1))
        kept = check(lib().qdb_gpu_rf_probe(self.f, hashes.ptr, selection.ptr if selection is not None else None, out.ptr,
                                            count), "rf_probe")
        return out, kept

    def stats(self) -> dict:
        v = (C.c_int64 * 5)()
        lib().qdb_gpu_rf_stats(self.f, v, 5)
        return {"kind": v[0], "exact_keys": v[1], "table_slots": v[2], "bloom_words": v[3], "bloom_bits": v[4]}

    def close(self):
        if self.f:
            lib().qdb_gpu_rf_free(self.f)
            self.f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GroupHash:
    """The group-hash entry point (rh_hash of the key columns of every
    selected row) for one schema and key count."""

    def __init__(self, types: Sequence[ValueType], nkeys: int):
        self.schema = schema_c(types)
        fn = C.c_void_p()
        self.slot = check(lib().qdb_gpu_compile_group_hash(C.byref(self.schema), nkeys, C.byref(fn)), "compile_group_hash")
        self.fn = abi.GROUP_HASH_FN(fn.value)
        self.nkeys = nkeys

    def __call__(self, rowset, key_cols: Sequence[int]) -> DeviceBuffer:
        out = DeviceBuffer(max(rowset.rows, 1) * 8)
        kc = (C.c_int64 * len(key_cols))(*key_cols)
        if not self.fn(rowset.ptr, out.ptr, kc):
            raise RuntimeError(f"group hash: {lib().qdb_gpu_last_error().decode()}")
        return out


class KernelRowSet(RowSetView):
    """A device rowset made by the library and owned here: released by
    close() (or when collected)."""

    def close(self):
        if self.rs is not None:
            lib().qdb_gpu_rowset_release(C.byref(self.rs))
            self.rs = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def shuffle_partition(rowset, outputs: int, routes: Optional[Sequence[int]] = None,
                      hashes: Optional[DeviceBuffer] = None):
    """The hash shuffle's routing of one batch (qdb_gpu_shuffle_partition):
    (row_index, offsets) with output o's rows, in input order, at
    row_index[offsets[o]:offsets[o + 1]] (row_index a DeviceBuffer of uint32,
    offsets a host int64 array)."""
    idx = DeviceBuffer(max(rowset.rows, 1) * 4)
    offs = np.zeros(outputs + 1, dtype=np.int64)
    r = None
    if routes is not None:
        r = np.ascontiguousarray(np.asarray(routes, dtype=np.int32))
    check(lib().qdb_gpu_shuffle_partition(rowset.ptr, hashes.ptr if hashes is not None else None, outputs,
                                          r.ctypes.data if r is not None else None, len(r) if r is not None else 0,
                                          idx.ptr, offs.ctypes.data_as(C.POINTER(C.c_int64))), "shuffle_partition")
    return idx, offs


def gather_rowset(rowset, row_index: DeviceBuffer, start: int, count: int) -> KernelRowSet:
    """Rows row_index[start:start + count] of a device batch as a new device
    batch (qdb_gpu_gather_rowset)."""
    schema = schema_c(rowset.types)
    out = abi.TRowSet()
    check(lib().qdb_gpu_gather_rowset(rowset.ptr, C.byref(schema), (row_index.ptr or 0) + 4 * start, count,
                                      C.byref(out)), "gather_rowset")
    return KernelRowSet(out, rowset.types)


class HashShuffle:
    """THashShuffleTask (qdb/scheduler/tasks/runtime_adapter.cpp:960-1030)
    over one batch on the device: each selected row goes to output
    rh_hash(keys) % outputs (or through a routing table), each output's rows
    in input order. The batch's own Hash column is used when it has one."""

    def __init__(self, types: Sequence[ValueType], key_cols: Sequence[int], outputs: int,
                 routes: Optional[Sequence[int]] = None):
        self.types = list(types)
        self.key_cols = list(key_cols)
        self.outputs = int(outputs)
        self.routes = routes
        self.hash = GroupHash(self.types, len(self.key_cols))

    def partition(self, rowset):
        hashes = None if rowset.rs.Hash else self.hash(rowset, self.key_cols)
        return shuffle_partition(rowset, self.outputs, self.routes, hashes)

    def __call__(self, rowset) -> List[KernelRowSet]:
        idx, offs = self.partition(rowset)
        return [gather_rowset(rowset, idx, int(offs[o]), int(offs[o + 1] - offs[o])) for o in range(self.outputs)]


def partial_states(kernels: AggregateKernels, batch) -> KernelRowSet:
    """A batch's single-row partial states (qdb_gpu_agg_partial_states)."""
    out = abi.TRowSet()
    check(lib().qdb_gpu_agg_partial_states(C.byref(kernels.k), batch.ptr, C.byref(out)), "agg_partial_states")
    lib().qdb_gpu_adopt_kernel_rowset(C.byref(out))
    return KernelRowSet(out, kernels.out_types)


class TPartialAggregateProcessor:
    """TPartialAggregateProcessor (qdb/exec/aggregate_exec.cpp:158-233) over
    the C-ABI: a bounded table that emits at max_groups groups and, with
    pass_through, passes rows through as partial states once sample_rows
    rows were hashed at fewer than min_reduction rows per group."""

    def __init__(self, kernels: AggregateKernels, initial_capacity: int = 8, sample_rows: int = 128 * 1024,
                 min_reduction: int = 8, max_groups: int = 1 << 20, pass_through: bool = True):
        self.kernels = kernels
        self.p = lib().qdb_gpu_partial_agg_new(C.byref(kernels.k), initial_capacity, sample_rows, min_reduction,
                                               max_groups, int(pass_through))
        if not self.p:
            raise RuntimeError(f"partial aggregate: {(lib().qdb_gpu_exec_last_error() or b'').decode()}")

    def add(self, batch) -> None:
        if lib().qdb_gpu_partial_agg_add(self.p, batch.ptr) != 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())

    def take(self) -> List[KernelRowSet]:
        """The outputs emitted so far."""
        outs = []
        while True:
            rs = abi.TRowSet()
            if not lib().qdb_gpu_partial_agg_next(self.p, C.byref(rs)):
                return outs
            outs.append(KernelRowSet(rs, self.kernels.out_types))

    def finish(self) -> Optional[KernelRowSet]:
        """The table's last groups; None when rows were passed through."""
        rs = abi.TRowSet()
        n = lib().qdb_gpu_partial_agg_finish(self.p, C.byref(rs))
        if n < 0:
            raise RuntimeError((lib().qdb_gpu_exec_last_error() or b"").decode())
        if not lib().qdb_gpu_partial_agg_passing_through(self.p):
            return KernelRowSet(rs, self.kernels.out_types)
        return None

    @property
    def passing_through(self) -> bool:
        return bool(lib().qdb_gpu_partial_agg_passing_through(self.p))

    def close(self):
        if self.p:
            lib().qdb_gpu_partial_agg_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SortKernel:
    """ORDER BY over a device batch (TSortProcessor's radix sort,
    qdb/exec/sort_exec.cpp:393-427): keys are (column, desc, nulls) with
    nulls one of abi.SORT_NULLS_*; the selected rows ordered, equal keys in
    input order."""

    def __init__(self, types: Sequence[ValueType], keys: Sequence[tuple]):
        self.types = list(types)
        self.schema = schema_c(self.types)
        self.nkeys = len(keys)
        self.keys = (abi.QdbSortKey * max(1, self.nkeys))()
        for i, k in enumerate(keys):
            col, desc, nulls = (list(k) + [False, abi.SORT_NULLS_DEFAULT])[:3]
            self.keys[i].Column = int(col)
            self.keys[i].Desc = int(bool(desc))
            self.keys[i].Nulls = int(nulls)

    def order(self, rowset):
        """(row_index DeviceBuffer of uint32, selected row count)."""
        idx = DeviceBuffer(max(rowset.rows, 1) * 4)
        cnt = C.c_int64(0)
        check(lib().qdb_gpu_sort_rowset(rowset.ptr, C.byref(self.schema), self.keys, self.nkeys, idx.ptr, C.byref(cnt)),
              "sort_rowset")
        return idx, int(cnt.value)

    def __call__(self, rowset, start: int = 0, limit: Optional[int] = None) -> KernelRowSet:
        """The ordered rows [start, start + limit) as a new device batch."""
        idx, n = self.order(rowset)
        start = max(0, min(start, n))
        count = n - start if limit is None else max(0, min(limit, n - start))
        return gather_rowset(rowset, idx, start, count)


def set_thread_stream(index: int) -> None:
    """The calling thread issues its library work on stream `index` (0: the
    library stream; up to 3 more)."""
    check(lib().qdb_gpu_set_thread_stream(index), "set_thread_stream")


def free_pair_buffer(pb) -> None:
    """Returns the device block of a PairBuffer filled by probe_detached."""
    if pb.Data:
        lib().qdb_gpu_free(pb.Data)
        pb.Data = None
        pb.Capacity = 0
        pb.Count = 0

