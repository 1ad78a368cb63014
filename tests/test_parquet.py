"""Parquet -> HBM: row-group pruning, decimal storage and the staged batches,
pinned by the reference's Parquet decimal goldens (its own decimals.parquet
file, test/regtest/cases/decimal_parquet/tables) and checked end to end
against the generator path."""
import decimal
import os

import numpy as np
import pytest

from helpers import DECIMAL, golden_rows, load_fixtures, parse_schema

pa = pytest.importorskip("pyarrow")
pq = pytest.importorskip("pyarrow.parquet")

HERE = os.path.dirname(os.path.abspath(__file__))
DECIMALS = os.path.join(HERE, "golden", "decimals.parquet")
FIX = load_fixtures()


def _pruning():
    """The scan module (its pruning and type mapping run on the host; no
    device call is made here)."""
    from qumirdb_amd import parquet
    return parquet


def _row_value(v, t):
    if v is None:
        return None
    if isinstance(v, decimal.Decimal):
        with decimal.localcontext() as ctx:
            ctx.prec = 100
            return int(v.scaleb(t.scale))
    return int(v)


def _eval(pred, row, P):
    """SQL three-valued evaluation of a pruning predicate on one row (dict
    of Python values), None for unknown."""
    if isinstance(pred, P.And):
        a, b = _eval(pred.left, row, P), _eval(pred.right, row, P)
        if a is False or b is False:
            return False
        return None if a is None or b is None else True
    if isinstance(pred, P.Or):
        a, b = _eval(pred.left, row, P), _eval(pred.right, row, P)
        if a is True or b is True:
            return True
        return None if a is None or b is None else False
    if isinstance(pred, P.IsNull):
        return row[pred.column] is None
    if isinstance(pred, P.IsNotNull):
        return row[pred.column] is not None
    v = row[pred.column]
    if v is None:
        return None
    x = decimal.Decimal(str(pred.value))
    v = decimal.Decimal(str(v))
    # Decimal comparisons are exact whatever the context's precision.
    return {"<": v < x, "<=": v <= x, ">": v > x, ">=": v >= x, "=": v == x, "!=": v != x}[pred.op]


PREDICATES = [
    ("d15 >= 2500007.57 and d38 > 0", lambda P: P.And(P.Cmp("d15", ">=", "2500007.57"), P.Cmp("d38", ">", 0))),
    ("id < 40", lambda P: P.Cmp("id", "<", 40)),
    ("id = 100 or id > 250", lambda P: P.Or(P.Cmp("id", "=", 100), P.Cmp("id", ">", 250))),
    ("d9 <= -100.5", lambda P: P.Cmp("d9", "<=", "-100.5")),
    ("d18 is null", lambda P: P.IsNull("d18")),
    ("d38 != 0", lambda P: P.Cmp("d38", "!=", 0)),
]


@pytest.mark.parametrize("text,make", PREDICATES)
def test_row_group_pruning_is_sound_and_follows_the_footer(text, make):
    """A pruned row group holds no row that satisfies the predicate, and the
    decision equals the interval test on the group's actual data (min, max
    and NULL count computed from its rows, not read from the footer)."""
    P = _pruning()
    pred = make(P)
    f = pq.ParquetFile(DECIMALS)
    cols = ["id", "d9", "d15", "d18", "d38"]
    types = {c: P.arrow_value_type(f.schema_arrow.field(c).type, True)[0] for c in cols}
    kept = 0
    for g in range(f.metadata.num_row_groups):
        rows = f.read_row_group(g, columns=cols).to_pylist()
        n = len(rows)
        footer = {}
        md = f.metadata.row_group(g)
        for i, c in enumerate(cols):
            st = md.column(i).statistics
            footer[c] = (st.min if st.has_min_max else None, st.max if st.has_min_max else None, st.null_count, types[c])
        data = {}
        for c in cols:
            vals = [r[c] for r in rows if r[c] is not None]
            data[c] = (min(vals) if vals else None, max(vals) if vals else None, n - len(vals), types[c])
        decision = P.may_be_true(pred, footer, n)
        assert decision == P.may_be_true(pred, data, n), g
        if not decision:
            assert not any(_eval(pred, r, P) is True for r in rows), g
        kept += decision
    assert 0 < kept <= f.metadata.num_row_groups


def test_filter_bounds_golden_from_pruned_scan():
    """decimal_parquet/filter_bounds: WHERE d15 >= 2500007.57 AND d38 > 0
    ORDER BY id over the reference's decimals.parquet, reading only the row
    groups the footer leaves (fewer than all)."""
    P = _pruning()
    pred = PREDICATES[0][1](P)
    f = pq.ParquetFile(DECIMALS)
    cols = ["id", "d9", "d15", "d38"]
    types = {c: P.arrow_value_type(f.schema_arrow.field(c).type, True)[0] for c in ["id", "d9", "d15", "d18", "d38"]}
    keep = []
    for g in range(f.metadata.num_row_groups):
        md = f.metadata.row_group(g)
        names = [md.column(i).path_in_schema for i in range(md.num_columns)]
        st = {c: (md.column(names.index(c)).statistics.min, md.column(names.index(c)).statistics.max,
                  md.column(names.index(c)).statistics.null_count, types[c]) for c in ["d15", "d38"]}
        if P.may_be_true(pred, st, md.num_rows):
            keep.append(g)
    assert 0 < len(keep) < f.metadata.num_row_groups
    rows = [r for g in keep for r in f.read_row_group(g, columns=cols).to_pylist() if _eval(pred, r, P) is True]
    rows.sort(key=lambda r: r["id"])
    fields = parse_schema(FIX["goldens"]["decimal_parquet/filter_bounds"]["schema"])
    got = [tuple(_row_value(r[c], t) for c, (_, t) in zip(cols, fields)) for r in rows]
    assert got == golden_rows(FIX["goldens"]["decimal_parquet/filter_bounds"])


def test_order_limit_golden_values_decode():
    """decimal_parquet/order_limit: the first four rows by id, every decimal
    width (FIXED_LEN_BYTE_ARRAY of 4, 7, 8 and 16 bytes) decoded to the
    golden's values."""
    t = pq.read_table(DECIMALS)
    rows = sorted(t.to_pylist(), key=lambda r: r["id"])[:4]
    fields = parse_schema(FIX["goldens"]["decimal_parquet/order_limit"]["schema"])
    got = [tuple(r[name] if t_.kind == 4 else _row_value(r[name], t_) for name, t_ in fields) for r in rows]
    assert got == golden_rows(FIX["goldens"]["decimal_parquet/order_limit"])


def test_storage_widths_follow_precision():
    """DECIMAL(p <= 18) is stored in 8 bytes, wider ones in 16
    (qdb/plan/types/decimal.h:80-99), whatever width Arrow decodes to."""
    P = _pruning()
    for p, width in [(1, 8), (9, 8), (15, 8), (18, 8), (19, 16), (38, 16)]:
        t, staged = P.arrow_value_type(pa.decimal128(p, 2), True)
        assert (t.kind, t.width, t.precision, staged) == (DECIMAL, width, p, 16)
    with pytest.raises(TypeError):
        P.arrow_value_type(pa.decimal256(40, 2), True)


# ---------------------------------------------------------------------------
# Device
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def gpu():
    import cases
    return cases.GpuBackend()


def _download(gpu, batch):
    return gpu.H.download_rowset(batch.rs, batch.types)


@pytest.mark.gpu
def test_staged_batches_equal_arrow(gpu):
    """Every row group of decimals.parquet staged into HBM: the decimals in
    their storage width (8, 8, 8 and 16 bytes) and the NULLs as Arrow
    decodes them."""
    from qumirdb_amd.parquet import ParquetSource
    cols = ["id", "d9", "d15", "d18", "d38"]
    src = ParquetSource(DECIMALS, columns=cols)
    assert [t.width for t in src.types] == [4, 8, 8, 8, 16]
    f = pq.ParquetFile(DECIMALS)
    for g, b in zip(src.row_groups, src.batches()):
        want = f.read_row_group(g, columns=cols).to_pylist()
        got = _download(gpu, b)
        for c, (name, t) in enumerate(zip(cols, src.types)):
            vals, valid = got[c]
            for i, r in enumerate(want):
                if r[name] is None:
                    assert valid is not None and not valid[i]
                else:
                    assert (valid is None or valid[i]) and int(vals[i]) == _row_value(r[name], t), (g, name, i)
        b.close()


@pytest.mark.gpu
def test_decimal_storage_kernel_widths(gpu):
    """qdb_gpu_decimal_storage from 4-, 8- and 16-byte sources to both
    storage widths: the value sign-extended, its low bytes kept."""
    H = gpu.H
    q = gpu.q
    rng = np.random.default_rng(4)
    n = 100_003
    v = rng.integers(-2 ** 31, 2 ** 31, n)
    for from_w, dt in [(4, np.int32), (8, np.int64)]:
        src = H.DeviceBuffer.from_numpy(v.astype(dt))
        for p, to in [(9, 8), (30, 16)]:
            dst = H.DeviceBuffer(n * to)
            q.check(q.lib().qdb_gpu_decimal_storage(src.ptr, from_w, n, p, dst.ptr), "storage")
            got = dst.to_numpy(np.int64, n * to // 8)
            if to == 8:
                assert np.array_equal(got, v)
            else:
                assert np.array_equal(got[0::2], v) and np.array_equal(got[1::2], np.where(v < 0, -1, 0))
    big = [int(x) * 10 ** 20 + int(y) for x, y in zip(rng.integers(-10 ** 6, 10 ** 6, n), rng.integers(0, 10 ** 6, n))]
    src16 = H.DeviceBuffer.from_numpy(H.i128_array(big))
    dst = H.DeviceBuffer(n * 16)
    q.check(q.lib().qdb_gpu_decimal_storage(src16.ptr, 16, n, 38, dst.ptr), "storage")
    assert H.i128_values(dst.to_numpy(np.dtype([("lo", "<u8"), ("hi", "<i8")]), n)) == big
    small = list(rng.integers(-10 ** 17, 10 ** 17, n))
    src16 = H.DeviceBuffer.from_numpy(H.i128_array(small))
    dst = H.DeviceBuffer(n * 8)
    q.check(q.lib().qdb_gpu_decimal_storage(src16.ptr, 16, n, 18, dst.ptr), "storage")
    assert dst.to_numpy(np.int64, n).tolist() == [int(x) for x in small]


def _one_batch(path, columns, predicate=None):
    from qumirdb_amd.parquet import ParquetSource
    src = ParquetSource(path, columns=columns, predicate=predicate, batch_rows=1 << 30)
    return src, src.read_batch()


@pytest.mark.gpu
def test_order_limit_golden_on_device(gpu):
    """decimal_parquet/order_limit on the device: ORDER BY id LIMIT 4 over
    the staged file (one batch of its 16 row groups)."""
    cols = ["id", "d9", "d15", "d18", "d38"]
    src, b = _one_batch(DECIMALS, cols)
    top = gpu.H.SortKernel(src.types, [(0, False, 0)])(b, start=0, limit=4)
    got = gpu.H.download_rowset(top.rs, src.types)
    rows = [tuple(None if (valid is not None and not valid[i]) else int(vals[i]) for vals, valid in got) for i in range(4)]
    want = [r[:5] for r in golden_rows(FIX["goldens"]["decimal_parquet/order_limit"])]
    assert rows == want
    top.close()
    b.close()


@pytest.mark.gpu
def test_filter_bounds_golden_on_device(gpu):
    """decimal_parquet/filter_bounds on the device: the pruned scan, the
    predicate as a filter kernel (decimal literals at the columns' scales),
    ORDER BY id."""
    from qumirdb_amd.parquet import And, Cmp
    H = gpu.H
    A = gpu.q.abi
    cols = ["id", "d9", "d15", "d38"]
    src, b = _one_batch(DECIMALS, cols, And(Cmp("d15", ">=", "2500007.57"), Cmp("d38", ">", 0)))
    assert 0 < len(src.row_groups) < 16
    p = H.Program(src.types)
    c1 = p.cmp(A.CMP_GE, p.col(2), p.const_decimal(250000757, 15, 2))
    c2 = p.cmp(A.CMP_GT, p.col(3), p.const_int(0))
    r = p.and_(c1, c2)
    sel = H.FilterKernel(src.types, p, r)(b)
    b.rs.Selection = sel.ptr
    out = H.SortKernel(src.types, [(0, False, 0)])(b)
    got = H.download_rowset(out.rs, src.types)
    rows = [tuple(None if (valid is not None and not valid[i]) else int(vals[i]) for vals, valid in got)
            for i in range(out.rows)]
    assert rows == golden_rows(FIX["goldens"]["decimal_parquet/filter_bounds"])
    out.close()
    b.close()


@pytest.mark.gpu
def test_layouts_agree_golden(gpu, tmp_path):
    """decimal_parquet/layouts_agree: DECIMAL(18,2) written as
    FIXED_LEN_BYTE_ARRAY and as INT64 stage to the same values, the
    golden's."""
    t = FIX["tables"]["decimal_parquet.layout"]
    ids = [r[0] for r in t["rows"]]
    vals = [None if r[1] is None else decimal.Decimal(r[1]) for r in t["rows"]]
    table = pa.table({"id": pa.array(ids, pa.int32()), "v": pa.array(vals, pa.decimal128(18, 2))})
    flba, i64 = str(tmp_path / "flba.parquet"), str(tmp_path / "i64.parquet")
    pq.write_table(table, flba)
    pq.write_table(table, i64, store_decimal_as_integer=True)
    assert pq.ParquetFile(i64).metadata.row_group(0).column(1).physical_type == "INT64"
    assert pq.ParquetFile(flba).metadata.row_group(0).column(1).physical_type == "FIXED_LEN_BYTE_ARRAY"
    outs = []
    for path in (flba, i64):
        src, b = _one_batch(path, ["id", "v"])
        got = _download(gpu, b)
        outs.append([tuple(None if (valid is not None and not valid[i]) else int(v_[i]) for v_, valid in got)
                     for i in range(b.rows)])
        b.close()
    assert outs[0] == outs[1]
    want = golden_rows(FIX["goldens"]["decimal_parquet/layouts_agree"])
    assert [(r[0], r[1]) for r in outs[0]] == [(w[0], w[1]) for w in want]


@pytest.mark.gpu
def test_lineitem_from_parquet_equals_generator(gpu, tmp_path):
    """The filtered lineitem group-by over 1M rows written to Parquet
    (DECIMAL(15,2) as FIXED_LEN_BYTE_ARRAY, row groups of 262,144 rows),
    staged batch by batch: the device result equals the one over the same
    rows generated in HBM."""
    from oracle import oracle as O
    from qumirdb_amd import workloads as W
    from qumirdb_amd.host import AggregateKernels, TAggregateProcessor
    from qumirdb_amd.parquet import ParquetSource
    n = 1_000_000
    cols = O.gen_lineitem(W.SEED_LINEITEM, 0, n)
    names = ["flag", "status", "shipdate", "quantity", "price", "discount", "tax"]
    arrays = [pa.array(cols[0], pa.int8()), pa.array(cols[1], pa.int8()),
              pa.array(cols[2].astype(np.int32), pa.int32()).cast(pa.date32())]
    for c in cols[3:]:
        # The unscaled int64 values as DECIMAL(15,2): Arrow's 16-byte
        # little-endian two's complement built directly.
        raw = np.empty((n, 2), np.int64)
        raw[:, 0] = c
        raw[:, 1] = np.where(c < 0, -1, 0)
        arrays.append(pa.Array.from_buffers(pa.decimal128(15, 2), n, [None, pa.py_buffer(raw.tobytes())]))
    schema = pa.schema([pa.field(nm, a.type, nullable=False) for nm, a in zip(names, arrays)])
    path = str(tmp_path / "lineitem.parquet")
    pq.write_table(pa.Table.from_arrays(arrays, schema=schema), path, row_group_size=262_144)
    src = ParquetSource(path)
    assert [t.width for t in src.types] == [1, 1, 4, 8, 8, 8, 8]
    ak = AggregateKernels(W.lineitem_fused_spec())
    proc = TAggregateProcessor(ak, initial_capacity=8)
    batches = list(src.batches())
    assert len(batches) == 4
    for b in batches:
        proc.add(b)
    got = W.lineitem_finalize_on_device(proc, ak)
    dev = W.LineitemDevice(n)
    proc2 = TAggregateProcessor(ak, initial_capacity=8)
    proc2.add(dev.rowset)
    want = W.lineitem_finalize_on_device(proc2, ak)
    assert got == want
    for b in batches:
        b.close()
