"""ORDER BY: the oracle's restatement of the reference's radix sort pinned by
the reference's ORDER BY goldens and by SQL ordering rules stated another
way (a comparison sort), and the device sort against the oracle."""
import functools
import math

import numpy as np
import pytest

import cases
from cases import Batch, GpuBackend, OracleBackend
from helpers import BOOL, DECIMAL, FLOAT, SIGNED, UNSIGNED, T, golden_rows, load_fixtures, np_column, parse_schema

from oracle import oracle as O

# (golden, ORDER BY as (column name, desc)) of the reference's regression
# cases (test/regtest/cases/<golden>.sql).
ORDERED_GOLDENS = [
    ("outer_join/null_keys", [("cv", False), ("ov", False)]),
    ("outer_join/duplicates", [("cv", False), ("ov", False)]),
    ("decimal_parquet/group_by", [("k", False), ("d2", False)]),
    ("decimal_parquet/join", [("id", False)]),
]


def _sql_cmp_key(types, columns, valid, keys):
    """SQL ORDER BY as a comparison: per key NULLs last ascending and first
    descending unless placed, floats with every NaN above +inf and -0 equal
    to +0, equal keys left in input order (Python's sort is stable)."""
    def value(col, row):
        if valid is not None and valid[col] is not None and not valid[col][row]:
            return None
        v = columns[col][row]
        t = types[col]
        if t.kind == FLOAT:
            v = float(v)
            return (1, 0.0) if math.isnan(v) else (0, v + 0.0)
        if t.kind == BOOL:
            return int(bool(v))
        return int(v)

    def cmp(a, b):
        for col, desc, nulls in keys:
            va, vb = value(col, a), value(col, b)
            first = desc if nulls == 0 else nulls == 1
            if va is None or vb is None:
                if va is None and vb is None:
                    continue
                a_first = (va is None) == first
                return -1 if a_first else 1
            if va == vb:
                continue
            less = va < vb
            return (-1 if less else 1) * (-1 if desc else 1)
        return 0
    return functools.cmp_to_key(cmp)


def _random_table(rng, n):
    types = [T(SIGNED, 4, nullable=True), T(FLOAT, 8), T(SIGNED, 1), T(DECIMAL, 16, 2, True, 30), T(BOOL, 1),
             T(UNSIGNED, 8), T(FLOAT, 8, nullable=True), T(SIGNED, 8)]
    f = rng.normal(0, 3, n).round(1)
    f[rng.random(n) < 0.05] = np.nan
    f[rng.random(n) < 0.05] = -0.0
    f[rng.random(n) < 0.05] = 0.0
    f[rng.random(n) < 0.02] = np.inf
    f[rng.random(n) < 0.02] = -np.inf
    g = rng.integers(-3, 3, n).astype(np.float64)
    g[rng.random(n) < 0.1] = np.nan
    cols = [rng.integers(-5, 5, n).astype(np.int32), f, rng.integers(-128, 128, n).astype(np.int8),
            [int(x) * 10 ** 25 + int(y) for x, y in zip(rng.integers(-3, 3, n), rng.integers(-2, 2, n))],
            rng.random(n) > 0.5, rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + np.uint64(rng.integers(0, 2)),
            g, rng.integers(-10 ** 18, 10 ** 18, n)]
    cols[5][:3] = [0, 2 ** 64 - 1, 2 ** 63]
    valid = [rng.random(n) > 0.2, None, None, rng.random(n) > 0.3, None, None, rng.random(n) > 0.25, None]
    return types, cols, valid


KEY_SETS = [
    [(0, False, 0)], [(0, True, 0)], [(0, False, 1)], [(0, True, 2)],
    [(1, False, 0), (2, True, 0)], [(3, False, 0), (0, True, 1)], [(3, True, 0)],
    [(4, False, 0), (6, True, 0), (1, False, 0)], [(5, False, 0)], [(5, True, 0), (7, False, 0)],
    [(6, False, 2), (0, False, 2), (2, False, 0)],
]


@pytest.mark.parametrize("keys", KEY_SETS)
def test_oracle_sort_equals_sql_ordering(keys):
    rng = np.random.default_rng(len(keys) * 7 + keys[0][0])
    n = 600
    types, cols, valid = _random_table(rng, n)
    sel = rng.random(n) > 0.15
    got = O.sort_rows(types, cols, valid, sel, keys)
    rows = [i for i in range(n) if sel[i]]
    want = sorted(rows, key=_sql_cmp_key(types, cols, valid, keys))
    assert got == want


def _golden_table(name):
    g = load_fixtures()["goldens"][name]
    fields = parse_schema(g["schema"])
    rows = golden_rows(g)
    return fields, rows


def _shuffled_columns(fields, rows, seed):
    perm = np.random.default_rng(seed).permutation(len(rows))
    shuffled = [rows[i] for i in perm]
    types = [t for _, t in fields]
    cols, valid = [], []
    for c, t in enumerate(types):
        vals = [r[c] for r in shuffled]
        v = np.array([x is not None for x in vals])
        if t.kind == FLOAT:
            vals = [0.0 if x is None else x for x in vals]
        elif t.kind == 4:  # strings are carried, not sorted on
            vals = [x or "" for x in vals]
        else:
            vals = [0 if x is None else x for x in vals]
        cols.append(vals)
        valid.append(v if t.nullable else None)
    return types, cols, valid, shuffled


def _assert_golden_order(fields, golden, ordered, keys):
    """The ordered rows equal the golden's row for row, up to the order of
    rows whose sort keys are equal (the reference leaves those in its input
    order, which a shuffled input does not reproduce)."""
    kc = [c for c, _, _ in keys]
    assert [tuple(r[c] for c in kc) for r in ordered] == [tuple(r[c] for c in kc) for r in golden]
    i = 0
    while i < len(golden):
        j = i
        while j < len(golden) and tuple(golden[j][c] for c in kc) == tuple(golden[i][c] for c in kc):
            j += 1
        assert sorted(map(repr, ordered[i:j])) == sorted(map(repr, golden[i:j]))
        i = j


@pytest.mark.parametrize("name,order", ORDERED_GOLDENS)
def test_oracle_sort_reproduces_reference_goldens(name, order):
    fields, rows = _golden_table(name)
    names = [f for f, _ in fields]
    keys = [(names.index(c), desc, 0) for c, desc in order]
    for seed in range(3):
        types, cols, valid, shuffled = _shuffled_columns(fields, rows, seed)
        idx = O.sort_rows(types, cols, valid, None, keys)
        _assert_golden_order(fields, rows, [shuffled[i] for i in idx], keys)


@pytest.mark.parametrize("wrong", [[("k", False, 1), ("d2", False, 0)], [("k", False, 0), ("d2", False, 1)],
                                   [("k", True, 2), ("d2", False, 0)]])
def test_oracle_sort_golden_check_catches_a_wrong_rule(wrong):
    """The golden check rejects NULLs placed first in an ascending key, and a
    descending key (decimal_parquet/group_by orders by nullable k and d2,
    NULLs last)."""
    fields, rows = _golden_table("decimal_parquet/group_by")
    names = [f for f, _ in fields]
    keys = [(names.index(c), d, nl) for c, d, nl in wrong]
    types, cols, valid, shuffled = _shuffled_columns(fields, rows, 0)
    idx = O.sort_rows(types, cols, valid, None, keys)
    with pytest.raises(AssertionError):
        _assert_golden_order(fields, rows, [shuffled[i] for i in idx], [(names.index("k"), False, 0),
                                                                          (names.index("d2"), False, 0)])


# ---------------------------------------------------------------------------
# Device
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def gpu():
    return GpuBackend()


def _device_order(gpu, types, cols, valid, sel, keys):
    H = gpu.H
    d = gpu.dev(Batch(types, cols, valid, None if sel is None else np.asarray(sel, np.uint8)))
    vts = [H.ValueType(t.kind, t.width, t.scale, t.nullable, t.precision) for t in types]
    sk = H.SortKernel(vts, keys)
    idx, n = sk.order(d)
    return idx.to_numpy(np.uint32, n) if n else np.zeros(0, np.uint32), d, sk


@pytest.mark.gpu
@pytest.mark.parametrize("keys", KEY_SETS)
def test_device_sort_equals_oracle(gpu, keys):
    rng = np.random.default_rng(len(keys) * 11 + keys[0][0])
    n = 20_000
    types, cols, valid = _random_table(rng, n)
    sel = rng.random(n) > 0.15
    got, _, _ = _device_order(gpu, types, cols, valid, sel, keys)
    assert got.tolist() == O.sort_rows(types, cols, valid, sel, keys)


@pytest.mark.gpu
@pytest.mark.parametrize("name,order", ORDERED_GOLDENS)
def test_device_sort_reproduces_reference_goldens(gpu, name, order):
    fields, rows = _golden_table(name)
    names = [f for f, _ in fields]
    keys = [(names.index(c), desc, 0) for c, desc in order]
    kept = [i for i, (_, t) in enumerate(fields) if t.kind != 4]
    for seed in range(2):
        types, cols, valid, shuffled = _shuffled_columns(fields, rows, seed)
        # String columns stay on the host: the device batch holds the others.
        sub_types = [types[i] for i in kept]
        sub_cols = [np_column(types[i], cols[i]) for i in kept]
        sub_valid = [valid[i] for i in kept]
        sub_keys = [(kept.index(c), d, nl) for c, d, nl in keys]
        idx, _, _ = _device_order(gpu, sub_types, sub_cols, sub_valid, None, sub_keys)
        _assert_golden_order(fields, rows, [shuffled[int(i)] for i in idx], keys)


@pytest.mark.gpu
def test_device_sort_top_slice_and_gather(gpu):
    """ORDER BY ... LIMIT: a slice of the order gathered into a batch holds
    the oracle's rows in the oracle's order, every column and NULL."""
    rng = np.random.default_rng(5)
    n = 50_000
    types, cols, valid = _random_table(rng, n)
    keys = [(1, True, 0), (0, False, 0)]
    order = O.sort_rows(types, cols, valid, None, keys)
    H = gpu.H
    d = gpu.dev(Batch(types, cols, valid))
    vts = [H.ValueType(t.kind, t.width, t.scale, t.nullable, t.precision) for t in types]
    top = H.SortKernel(vts, keys)(d, start=100, limit=1000)
    got = H.download_rowset(top.rs, vts)
    host = OracleBackend().host(Batch(types, cols, valid))
    want = O.read_host_rowset(host.rs, types)
    rows = order[100:1100]
    for c, ((gv, gm), (wv, wm)) in enumerate(zip(got, want)):
        assert [x for x in gv] == [wv[r] for r in rows] or all(
            (isinstance(a, float) and math.isnan(a) and math.isnan(b)) or a == b for a, b in zip(gv, [wv[r] for r in rows])), c
        if wm is not None:
            assert np.array_equal(gm, np.asarray(wm)[rows]), c
    top.close()


@pytest.mark.gpu
def test_device_sort_full_size_properties(gpu):
    """20M rows by a Zipf-like i64 key descending then a row-unique i64
    ascending: a permutation of the rows whose keys are non-increasing, and
    within equal first keys the second key increases."""
    rng = np.random.default_rng(9)
    n = 20_000_000
    a = (rng.zipf(1.3, n) % 100_000).astype(np.int64)
    b = rng.permutation(n).astype(np.int64)
    types = [T(SIGNED, 8), T(SIGNED, 8)]
    got, _, _ = _device_order(gpu, types, [a, b], None, None, [(0, True, 0), (1, False, 0)])
    got = np.asarray(got, dtype=np.int64)
    assert len(got) == n and np.array_equal(np.sort(got), np.arange(n))
    ka, kb = a[got], b[got]
    assert np.all(np.diff(ka) <= 0)
    same = np.diff(ka) == 0
    assert np.all(np.diff(kb)[same] > 0)


def _order_wide_input():
    """decimal_parquet/order_wide: SELECT id, d2, d10 FROM w WHERE d10 IS NOT
    NULL ORDER BY d10 DESC, id LIMIT 4 -- the wide table with its selection."""
    from helpers import table_columns
    names, types, values, valid = table_columns(load_fixtures()["tables"]["decimal_parquet.wide"])
    sel = valid[names.index("d10")].copy()
    keys = [(names.index("d10"), True, 0), (names.index("id"), False, 0)]
    return names, types, values, valid, sel, keys


def _order_wide_rows(names, values, valid, order):
    cols = [names.index(c) for c in ("id", "d2", "d10")]
    return [tuple(None if (valid[c] is not None and not valid[c][r]) else int(values[c][r]) for c in cols) for r in order[:4]]


def test_oracle_sort_order_wide_golden():
    names, types, values, valid, sel, keys = _order_wide_input()
    order = O.sort_rows(types, values, valid, sel, keys)
    assert _order_wide_rows(names, values, valid, order) == golden_rows(load_fixtures()["goldens"]["decimal_parquet/order_wide"])


@pytest.mark.gpu
def test_device_sort_order_wide_golden(gpu):
    """DESC on a 38-digit decimal, then id, with a selection, LIMIT 4."""
    names, types, values, valid, sel, keys = _order_wide_input()
    cols = [np_column(t, v) for t, v in zip(types, values)]
    order, _, _ = _device_order(gpu, types, cols, valid, sel, keys)
    assert _order_wide_rows(names, values, valid, order.tolist()) == \
        golden_rows(load_fixtures()["goldens"]["decimal_parquet/order_wide"])
