"""Parquet -> HBM: the scan side of the path.

The reference reads Parquet with Arrow into batches of at most 262,144 rows
(TParquetFileData::Open, qdb/io/parquet/source.cpp:320-348), hands Arrow's
buffers to the kernels without a copy (RecordBatchToRowSet, :98-190),
stores DECIMAL(p <= 18) as int64 and wider ones as int128
(DecimalStorageValues, qdb/io/parquet/decimal_storage.cpp:36-60) and skips
row groups whose footer statistics rule the scan's predicate out
(TRowGroupPredicateEvaluator, qdb/io/parquet/row_group_predicate.cpp).

Here pyarrow decodes on the host (the reference's Arrow); each batch's
buffers are staged into HBM with one pinned copy (qdb_gpu_stage_rowset),
and decimal columns are converted to their storage width on the device
(qdb_gpu_decimal_storage). Row groups are pruned with the same kind of
interval test over the footer's min / max / null counts. String columns are
not supported on the device and cannot be selected.
"""
from __future__ import annotations

import ctypes as C
import decimal
from dataclasses import dataclass
from typing import List, Optional, Sequence, Union

import numpy as np

from . import abi, check, lib
from .host import DeviceBuffer, KernelRowSet, ValueType, schema_c

BATCH_ROWS = 1 << 18  # source.cpp:333


# ---------------------------------------------------------------------------
# Types
# ---------------------------------------------------------------------------

def arrow_value_type(t, nullable: bool) -> tuple[ValueType, int]:
    """(the kernels' type, the staged width in bytes) of an Arrow type."""
    import pyarrow as pa
    if pa.types.is_boolean(t):
        return ValueType(abi.BOOL, 1, nullable=nullable), 1
    if pa.types.is_integer(t):
        kind = abi.SIGNED if pa.types.is_signed_integer(t) else abi.UNSIGNED
        return ValueType(kind, t.bit_width // 8, nullable=nullable), t.bit_width // 8
    if pa.types.is_floating(t):
        if t.bit_width == 16:
            raise TypeError("half floats are not supported")
        return ValueType(abi.FLOAT, t.bit_width // 8, nullable=nullable), t.bit_width // 8
    if pa.types.is_date32(t):
        return ValueType(abi.DATE, 4, nullable=nullable), 4
    if pa.types.is_decimal(t):
        if t.precision > 38:
            raise TypeError(f"DECIMAL({t.precision}) has no storage")
        storage = 8 if t.precision <= 18 else 16
        return ValueType(abi.DECIMAL, storage, t.scale, nullable=nullable, precision=t.precision), t.byte_width
    raise TypeError(f"Arrow type {t} is not supported on the device")


# ---------------------------------------------------------------------------
# Row-group pruning
# ---------------------------------------------------------------------------

@dataclass
class Cmp:
    """column <op> literal, op one of < <= > >= = !=."""
    column: str
    op: str
    value: Union[int, float, str, decimal.Decimal]


@dataclass
class And:
    left: object
    right: object


@dataclass
class Or:
    left: object
    right: object


@dataclass
class IsNull:
    column: str


@dataclass
class IsNotNull:
    column: str


def _unscaled(v, scale: int) -> int:
    """An exact value at `scale` decimal places (a literal or a statistic);
    38-digit decimals need more than the default 28 digits of context."""
    with decimal.localcontext() as ctx:
        ctx.prec = 100
        d = decimal.Decimal(str(v)) if not isinstance(v, decimal.Decimal) else v
        q = d.scaleb(scale)
        if q != q.to_integral_value():
            raise ValueError("not exact")
        return int(q)


def _literal_scale(v) -> int:
    if isinstance(v, (int, np.integer)):
        return 0
    d = decimal.Decimal(str(v)) if not isinstance(v, decimal.Decimal) else v
    return max(0, -d.as_tuple().exponent)


def may_be_true(pred, stats: dict, num_rows: int) -> bool:
    """Whether a row of a row group can satisfy `pred` given its columns'
    statistics ({column: (min, max, null_count, type)} with min/max None when
    absent): the interval test of the reference's row-group predicate, which
    answers "maybe" for anything it cannot decide."""
    if isinstance(pred, And):
        return may_be_true(pred.left, stats, num_rows) and may_be_true(pred.right, stats, num_rows)
    if isinstance(pred, Or):
        return may_be_true(pred.left, stats, num_rows) or may_be_true(pred.right, stats, num_rows)
    if isinstance(pred, (IsNull, IsNotNull)):
        st = stats.get(pred.column)
        if st is None or st[2] is None:
            return True
        nulls = st[2]
        return nulls > 0 if isinstance(pred, IsNull) else nulls < num_rows
    if isinstance(pred, Cmp):
        st = stats.get(pred.column)
        if st is None:
            return True
        lo, hi, nulls, t = st
        if nulls is not None and nulls >= num_rows:
            return False  # a comparison with NULL is never true
        if lo is None or hi is None:
            return True
        try:
            if t.kind == abi.DECIMAL or isinstance(pred.value, (str, decimal.Decimal)):
                scale = max(t.scale if t.kind == abi.DECIMAL else 0, _literal_scale(pred.value))
                lo_v, hi_v, x = _unscaled(lo, scale), _unscaled(hi, scale), _unscaled(pred.value, scale)
            elif t.kind == abi.FLOAT:
                lo_v, hi_v, x = float(lo), float(hi), float(pred.value)
                if x != x:
                    return pred.op == "!="
            else:
                lo_v, hi_v = int(lo), int(hi)
                x = pred.value
                if isinstance(x, float) and x != int(x):
                    lo_v, hi_v = float(lo_v), float(hi_v)
                else:
                    x = int(x)
        except (ValueError, decimal.InvalidOperation):
            return True
        op = pred.op
        if op == "<":
            return lo_v < x
        if op == "<=":
            return lo_v <= x
        if op == ">":
            return hi_v > x
        if op == ">=":
            return hi_v >= x
        if op == "=":
            return lo_v <= x <= hi_v
        if op == "!=":
            return not (lo_v == hi_v == x)
        raise ValueError(f"unknown comparison {op}")
    raise TypeError(f"unknown predicate node {pred!r}")


# ---------------------------------------------------------------------------
# Device batches
# ---------------------------------------------------------------------------

class ParquetBatch:
    """A device batch staged from Parquet: kernel input like any device
    batch (.rs / .ptr / .types / .rows); holds its device buffers."""

    def __init__(self, types: Sequence[ValueType], staged: KernelRowSet, converted: dict):
        self.types = list(types)
        self.rows = int(staged.rs.RowCount)
        self._staged = staged
        self._converted = converted
        n = len(self.types)
        self.cols = (abi.TColumn * max(1, n))()
        for c in range(n):
            self.cols[c] = staged.rs.Columns[c]
            if c in converted:
                self.cols[c].Data = converted[c].ptr
        self.rs = abi.TRowSet()
        self.rs.Columns = C.cast(self.cols, C.POINTER(abi.TColumn))
        self.rs.ColumnCount = n
        self.rs.RowCount = self.rows
        self.rs.RefCount = 1

    @property
    def ptr(self):
        return C.byref(self.rs)

    def close(self):
        self._converted = {}
        if self._staged is not None:
            self._staged.close()
            self._staged = None


class ParquetSource:
    """Batches of a Parquet file's columns in HBM (TParquetSource,
    qdb/io/parquet/source.cpp): the row groups the predicate may match, in
    file order, at most `batch_rows` rows per batch."""

    def __init__(self, path: str, columns: Optional[Sequence[str]] = None, predicate=None, batch_rows: int = BATCH_ROWS):
        import pyarrow.parquet as pq
        self.file = pq.ParquetFile(path)
        arrow_schema = self.file.schema_arrow
        self.columns = list(columns) if columns is not None else list(arrow_schema.names)
        self.fields = [arrow_schema.field(c) for c in self.columns]
        typed = [arrow_value_type(f.type, f.nullable) for f in self.fields]
        self.types = [t for t, _ in typed]
        self.staged_widths = [w for _, w in typed]
        self.batch_rows = int(batch_rows)
        self.predicate = predicate
        md = self.file.metadata
        self.row_groups = [g for g in range(md.num_row_groups) if predicate is None or may_be_true(predicate, self.stats(g),
                                                                                                    md.row_group(g).num_rows)]

    def stats(self, g: int) -> dict:
        """{column: (min, max, null_count, type)} of row group g's footer."""
        md = self.file.metadata.row_group(g)
        names = [md.column(i).path_in_schema for i in range(md.num_columns)]
        out = {}
        for name, t in zip(self.columns, self.types):
            if name not in names:
                continue
            st = md.column(names.index(name)).statistics
            if st is None:
                out[name] = (None, None, None, t)
                continue
            lo = st.min if st.has_min_max else None
            hi = st.max if st.has_min_max else None
            out[name] = (lo, hi, st.null_count if st.has_null_count else None, t)
        return out

    def _stage(self, table) -> ParquetBatch:
        n = table.num_rows
        arrays = [table.column(c).combine_chunks() for c in self.columns]
        cols = (abi.TColumn * max(1, len(arrays)))()
        keep = []
        for c, (arr, t, w) in enumerate(zip(arrays, self.types, self.staged_widths)):
            bufs = arr.buffers()
            validity, values = bufs[0], bufs[1]
            keep.append(arr)
            if t.kind == abi.BOOL:
                cols[c].Data = values.address
                cols[c].DataBitOffset = arr.offset
            else:
                cols[c].Data = values.address + arr.offset * w
            if validity is not None and arr.null_count > 0:
                cols[c].Mask = validity.address
                cols[c].MaskBitOffset = arr.offset
        host = abi.TRowSet()
        host.Columns = C.cast(cols, C.POINTER(abi.TColumn))
        host.ColumnCount = len(arrays)
        host.RowCount = n
        host.RefCount = 1
        # Staged as Arrow lays the values out (decimals at Arrow's width).
        stage_types = [ValueType(t.kind, w if t.kind == abi.DECIMAL else t.width, t.scale, t.nullable, t.precision)
                       for t, w in zip(self.types, self.staged_widths)]
        schema = schema_c(stage_types)
        out = abi.TRowSet()
        check(lib().qdb_gpu_stage_rowset(C.byref(host), C.byref(schema), C.byref(out)), "stage_rowset")
        staged = KernelRowSet(out, stage_types)
        converted = {}
        for c, (t, w) in enumerate(zip(self.types, self.staged_widths)):
            if t.kind == abi.DECIMAL and w != t.width:
                buf = DeviceBuffer(max(n, 1) * t.width)
                check(lib().qdb_gpu_decimal_storage(out.Columns[c].Data, w, n, t.precision, buf.ptr), "decimal_storage")
                converted[c] = buf
        del keep
        return ParquetBatch(self.types, staged, converted)

    def batches(self):
        """Device batches of the surviving row groups, in file order."""
        for g in self.row_groups:
            table = self.file.read_row_group(g, columns=self.columns)
            for start in range(0, max(table.num_rows, 1), self.batch_rows):
                part = table.slice(start, self.batch_rows)
                if part.num_rows == 0 and table.num_rows > 0:
                    continue
                yield self._stage(part)

    def read_all(self) -> List[ParquetBatch]:
        return list(self.batches())

    def read_batch(self) -> ParquetBatch:
        """Every surviving row group as one device batch (an operator that
        needs all rows at once, such as ORDER BY over a small table)."""
        table = self.file.read_row_groups(self.row_groups, columns=self.columns)
        if table.num_rows >= (1 << 31):
            raise ValueError("a batch holds fewer than 2^31 rows")
        return self._stage(table)
