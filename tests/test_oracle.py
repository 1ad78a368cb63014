"""Pins the CPU oracle: the reference's own golden results for the path,
closed forms (exact rational rounding), and brute force on small inputs."""
import math
import random
from fractions import Fraction

import numpy as np
import pytest

import cases
from cases import AGG_COUNT, AGG_COUNT_STAR, AGG_MAX, AGG_MIN, AGG_SUM, AggCase, Batch, JoinCase, OracleBackend
from helpers import DECIMAL, FLOAT, SIGNED, T, golden_rows, result_rows, sort_key

from oracle import oracle as O


@pytest.fixture(scope="module")
def backend():
    return OracleBackend(fsum_mode=0)


def same(got, want):
    assert len(got) == len(want), f"{len(got)} rows vs {len(want)}\n{got}\n{want}"
    for g, w in zip(got, want):
        assert len(g) == len(w)
        for a, b in zip(g, w):
            if isinstance(b, float) or isinstance(a, float):
                assert a is not None and b is not None and (a == b or math.isclose(a, b, rel_tol=0, abs_tol=0)), f"{g} vs {w}"
            else:
                assert a == b, f"{g} vs {w}"


@pytest.mark.parametrize("name", sorted(cases.AGG_GOLDENS))
def test_aggregate_goldens(backend, name):
    got, want = cases.AGG_GOLDENS[name](backend)
    same(got, want)


@pytest.mark.parametrize("name", sorted(cases.JOIN_GOLDENS))
def test_join_goldens(backend, name):
    got, want = cases.JOIN_GOLDENS[name](backend)
    same(got, want)


@pytest.mark.parametrize("name", sorted(cases.PROJECT_GOLDENS))
def test_project_goldens(backend, name):
    got, want = cases.PROJECT_GOLDENS[name](backend)
    same(got, want)


def test_decimal_casts_golden():
    """decimal_parquet/casts: CAST(d2 AS DOUBLE), CAST(d10 AS DOUBLE) are the
    correctly rounded doubles (qdb_decimal_to_f64)."""
    names, b = cases.table_batch("decimal_parquet.wide")
    d2, d10 = names.index("d2"), names.index("d10")
    golden = golden_rows(cases.FIX["goldens"]["decimal_parquet/casts"])
    by_id = {int(i): r for r, i in enumerate(b.columns[names.index("id")])}
    for gid, f2, f10, _ in golden:
        r = by_id[gid]
        if f2 is not None:
            assert O.decimal_to_f64(b.columns[d2][r], 2) == f2
        if f10 is not None:
            assert O.decimal_to_f64(b.columns[d10][r], 10) == f10


def nearest(fr: Fraction) -> float:
    return float(fr)  # CPython rounds a rational to the nearest double, ties to even


@pytest.mark.parametrize("scale", [0, 2, 4, 6, 10, 18, 22, 24, 30, 38])
def test_decimal_to_f64_exact_rounding(scale):
    rng = random.Random(scale)
    vals = [0, 1, -1, (1 << 53) - 1, 1 << 53, (1 << 53) + 1, -(1 << 53) - 1, 10 ** 38 - 1, -(10 ** 38 - 1),
            (1 << 63), 5 * 10 ** (scale - 1) if scale else 5, 12345678901234567890123456789012345678 % (10 ** 38)]
    vals += [rng.randrange(-(10 ** 38) + 1, 10 ** 38) for _ in range(200)]
    vals += [rng.randrange(-(1 << 60), 1 << 60) for _ in range(200)]
    for v in vals:
        assert O.decimal_to_f64(v, scale) == nearest(Fraction(v, 10 ** scale)), (v, scale)


def test_decimal_to_f64_ties_to_even():
    # 2^53 + 1 over 1: exactly between two doubles -> even mantissa.
    assert O.decimal_to_f64((1 << 53) + 1, 0) == float(1 << 53)
    assert O.decimal_to_f64((1 << 53) + 3, 0) == float((1 << 53) + 4)
    assert O.decimal_to_f64(((1 << 53) + 1) * 100, 2) == float(1 << 53)


def exact_fraction_sum(vals):
    return sum((Fraction(v) for v in vals), Fraction(0))


@pytest.mark.parametrize("seed", range(6))
def test_exact_sum_is_correctly_rounded(seed):
    rng = np.random.default_rng(seed)
    n = 1 + seed * 37
    vals = rng.normal(0, 1e3, n)
    if seed % 2:
        vals = np.concatenate([vals, -vals[:-1], [1e-300, 1e300, -1e300]])
    if seed == 4:
        vals = np.array([1e16, 1.0, -1e16, 1e-20, 3.0 * 2 ** -1074])
    want = float(exact_fraction_sum(vals.tolist()))
    assert O.exact_sum(vals) == want


def test_exact_sum_specials():
    assert math.isnan(O.exact_sum([1.0, float("nan")]))
    assert O.exact_sum([1.0, float("inf")]) == float("inf")
    assert math.isnan(O.exact_sum([float("inf"), float("-inf")]))
    assert O.exact_sum([]) == 0.0


def test_kahan_close_to_exact():
    rng = np.random.default_rng(7)
    vals = rng.normal(0, 1e3, 100000)
    assert abs(O.kahan_sum(vals) - O.exact_sum(vals)) <= 1e-12 * abs(O.exact_sum(vals))


def brute_groupby(keys, vals, valid):
    out = {}
    for k, v, ok in zip(keys, vals, valid):
        g = out.setdefault(k, {"n": 0, "cnt": 0, "sum": 0, "min": None, "max": None, "f": []})
        g["n"] += 1
        if ok:
            g["cnt"] += 1
            g["sum"] += v
            g["min"] = v if g["min"] is None else min(g["min"], v)
            g["max"] = v if g["max"] is None else max(g["max"], v)
    return out


@pytest.mark.parametrize("seed", range(4))
def test_aggregate_brute_force(seed):
    """Random nullable keys and values: counts, i64 wrap sums, i128 decimal
    sums, min/max and exact f64 sums against dict-based brute force."""
    rng = np.random.default_rng(seed)
    n = 5000
    k1 = rng.integers(-3, 4, n).astype(np.int32)
    k1v = rng.random(n) > 0.1
    k2 = rng.integers(0, 50, n).astype(np.int64)
    vi = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)
    viv = rng.random(n) > 0.2
    dec = [int(x) * (10 ** 20) + int(y) for x, y in zip(rng.integers(-10 ** 9, 10 ** 9, n), rng.integers(0, 10 ** 9, n))]
    decv = rng.random(n) > 0.3
    vf = rng.normal(0, 1e3, n)
    types = [T(SIGNED, 4, nullable=True), T(SIGNED, 8), T(SIGNED, 8, nullable=True), T(DECIMAL, 16, 3, True, 38), T(FLOAT, 8)]
    case = AggCase(types, [0, 1], [(AGG_COUNT_STAR, -1), (AGG_COUNT, 2), (AGG_SUM, 2), (AGG_MIN, 2), (AGG_MAX, 2),
                                   (AGG_SUM, 3), (AGG_MIN, 3), (AGG_MAX, 3), (AGG_SUM, 4)])
    b = Batch(types, [k1, k2, vi, dec, vf], [k1v, None, viv, decv, None])
    res = result_rows(OracleBackend(fsum_mode=2).aggregate(case, [b]))
    keys = [(int(a) if ok else None, int(c)) for a, ok, c in zip(k1, k1v, k2)]
    bi = brute_groupby(keys, [int(x) for x in vi], viv)
    bd = brute_groupby(keys, dec, decv)
    bf = {}
    for kk, f in zip(keys, vf):
        bf.setdefault(kk, []).append(f)
    assert len(res) == len(bi)
    wrap = lambda x: ((x + (1 << 63)) % (1 << 64)) - (1 << 63)
    for row in res:
        kk = (row[0], row[1])
        gi, gd = bi[kk], bd[kk]
        assert row[2] == gi["n"]
        assert row[3] == gi["cnt"]
        assert row[4] == (wrap(gi["sum"]) if gi["cnt"] else None)
        assert row[5] == gi["min"] and row[6] == gi["max"]
        assert row[7] == (gd["sum"] if gd["cnt"] else None)
        assert row[8] == gd["min"] and row[9] == gd["max"]
        assert row[10] == float(exact_fraction_sum(bf[kk]))


def test_aggregate_selection_and_batches():
    """Rows with Selection[i] == 0 are skipped; batches accumulate."""
    types = [T(SIGNED, 8), T(SIGNED, 8)]
    case = AggCase(types, [0], [(AGG_COUNT_STAR, -1), (AGG_SUM, 1)])
    b1 = Batch(types, [[1, 2, 1, 3], [10, 20, 30, 40]], selection=np.array([1, 1, 0, 1], np.uint8))
    b2 = Batch(types, [[1, 3], [5, 6]])
    got = sorted(result_rows(OracleBackend().aggregate(case, [b1, b2])), key=sort_key)
    assert got == [(1, 2, 15), (2, 1, 20), (3, 2, 46)]


def brute_join(bkeys, bvalid, pkeys, pvalid, psel, outer):
    pairs = []
    for i, (pk, pv) in enumerate(zip(pkeys, pvalid)):
        if psel is not None and not psel[i]:
            continue
        m = [j for j, (bk, bv) in enumerate(zip(bkeys, bvalid)) if pv and bv and bk == pk]
        for j in m:
            pairs.append((i, j))
        if not m and outer:
            pairs.append((i, -1))
    return sorted(pairs)


@pytest.mark.parametrize("outer", [False, True])
@pytest.mark.parametrize("seed", range(3))
def test_join_brute_force(seed, outer):
    rng = np.random.default_rng(seed)
    nb, npr = 300, 2000
    bk = rng.integers(0, 200, nb).astype(np.int64)
    bv = rng.random(nb) > 0.1
    bsel = rng.random(nb) > 0.3
    pk = rng.integers(0, 260, npr).astype(np.int64)
    pv = rng.random(npr) > 0.1
    psel = (rng.random(npr) > 0.2).astype(np.uint8)
    types = [T(SIGNED, 8, nullable=True)]
    case = JoinCase(types, types, [0], [0], join_type=cases.JOIN_LEFT_OUTER if outer else cases.JOIN_INNER)
    j = O.OracleJoin(case.to_c())
    j.append(O.HostRowSet(types, [bk], [bv], selection=bsel.astype(np.uint8)))
    j.build()
    pairs = j.probe(O.HostRowSet(types, [pk], [pv], selection=psel))
    got = sorted((int(a) + (1 << 32), int(b)) for a, b in pairs)
    want = brute_join(bk, bv & bsel, pk, pv, psel, outer)
    assert got == want


def test_join_chains_in_build_order():
    """Equal build keys are emitted in build order (jr_build_rows links the
    rows from the last to the first)."""
    types = [T(SIGNED, 8)]
    case = JoinCase(types, types, [0], [0])
    j = O.OracleJoin(case.to_c())
    j.append(O.HostRowSet(types, [np.array([5, 7, 5, 5, 9], np.int64)]))
    j.build()
    pairs = j.probe(O.HostRowSet(types, [np.array([5, 9], np.int64)]))
    assert [int(b) for _, b in pairs] == [0, 2, 3, 4]
    assert j.keys() == 3 and j.rows() == 5


@pytest.mark.parametrize("outer", [False, True])
def test_join_parallel_probe_equals_serial(outer):
    """The morsel-parallel probe (the CPU baseline's multi-thread leg) emits
    the serial probe's pairs in the same order: 700K probe rows over three
    morsels, duplicate build keys, NULL keys and a selection."""
    rng = np.random.default_rng(11)
    nb, npr = 5000, 700_000
    types = [T(SIGNED, 8, nullable=True)]
    case = JoinCase(types, types, [0], [0], join_type=cases.JOIN_LEFT_OUTER if outer else cases.JOIN_INNER)
    j = O.OracleJoin(case.to_c())
    j.append(O.HostRowSet(types, [rng.integers(0, 4000, nb).astype(np.int64)], [rng.random(nb) > 0.05]))
    j.build()
    probe = O.HostRowSet(types, [rng.integers(0, 5000, npr).astype(np.int64)], [rng.random(npr) > 0.05],
                         selection=(rng.random(npr) > 0.1).astype(np.uint8))
    serial = j.probe(probe)
    n = O.LIB.oracle_join_probe_parallel(j.h, probe.ptr, 4)
    assert n == len(serial) > 0
    par = np.ctypeslib.as_array(O.LIB.oracle_join_pairs(j.h), shape=(2 * n,)).reshape(n, 2)
    assert np.array_equal(par, serial)


def test_three_valued_filter():
    """NULL is not TRUE: x > 0 OR y > 0 with NULLs (SQL three-valued logic)."""
    E = cases._abi()
    types = [T(SIGNED, 8, nullable=True), T(SIGNED, 8, nullable=True)]
    x = [1, -1, 0, 1, 0, -1]
    xv = np.array([1, 1, 0, 0, 0, 1], bool)
    y = [0, 0, 1, 0, 0, 5]
    yv = np.array([0, 0, 1, 0, 0, 1], bool)
    p = cases.Prog(types)
    zero = p.const(0, T(SIGNED, 8))
    a = p.cmp(E.CMP_GT, p.col(0), zero)
    b = p.cmp(E.CMP_GT, p.col(1), zero)
    r = p._emit(E.EXPR_OR, E.REG_BOOL, T(3, 1, nullable=True), a=a, b=b)
    sel = OracleBackend().filter(types, p, r, Batch(types, [x, y], [xv, yv]))
    # row: x>0, y>0 -> OR
    # 0: T, NULL -> T ; 1: F, NULL -> NULL ; 2: NULL, F(1>0 T) -> T ; 3: NULL, NULL -> NULL
    # 4: NULL, NULL -> NULL ; 5: F, T -> T
    assert sel.tolist() == [1, 0, 1, 0, 0, 1]
    sel2 = OracleBackend().filter(types, p, r, Batch(types, [x, y], [xv, yv], selection=np.array([0, 1, 1, 1, 1, 1], np.uint8)))
    assert sel2.tolist() == [0, 0, 1, 0, 0, 1]


def test_zipf_cdf_host_matches_numpy():
    """The host generator's CDF equals the numpy cumulative sum the device
    workloads upload (qumirdb_amd/workloads.py zipf1_cdf)."""
    n = 10000
    h = np.cumsum(1.0 / np.arange(1, n + 1, dtype=np.float64))
    cdf = h / h[-1]
    cdf[-1] = 1.0
    assert np.array_equal(O.zipf_cdf(n), cdf)


def test_lineitem_generator_ranges():
    flag, status, ship, qty, price, disc, tax = O.gen_lineitem(0xC0, 0, 100000)
    assert set(np.unique(flag)) == {0, 1, 2} and set(np.unique(status)) == {0, 1}
    assert ship.min() >= 0 and ship.max() <= 2525
    assert set(np.unique(qty // 100)) == set(range(1, 51)) and (qty % 100 == 0).all()
    assert price.min() >= 90000 and price.max() <= 10500000
    assert disc.min() == 0 and disc.max() == 10 and tax.min() == 0 and tax.max() == 8
    sel = (ship <= 2436).mean()
    assert abs(sel - 2437 / 2526) < 0.005
    # counter-based: a slice regenerates the same rows
    again = O.gen_lineitem(0xC0, 500, 100)
    assert np.array_equal(again[4], price[500:600])


def test_join_generator_permutation():
    key, seg = O.gen_join_build(0xC1, 0, 100000, 100000)
    assert np.array_equal(np.sort(key), np.arange(1, 100001))
    assert abs((seg == 1).mean() - 0.2) < 0.01
    fk, val = O.gen_join_probe(0xC1, 0, 100000, 100000)
    assert fk.min() >= 1 and fk.max() <= 100000


def test_groups_generator_zipf():
    cdf = O.zipf_cdf(10000)
    key, vi, vf = O.gen_groups(0xC2, 0, 200000, cdf)
    vals, counts = np.unique(key, return_counts=True)
    top = counts.max() / len(key)
    h = np.sum(1.0 / np.arange(1, 10001))
    assert abs(top - 1.0 / h) < 0.01
    assert vi.min() >= -1000000 and vi.max() <= 1000000
    assert abs(vf.std() - 1000.0) < 20.0


# ---------------------------------------------------------------------------
# rh_hash: known answers worked out by hand from qdb/kernel/gen.cpp:233-379
# (HashKeyValue) over the key struct of qdb/kernel/aggregate_key.cpp:246-272.
# ---------------------------------------------------------------------------

def _key_hash(types, cols, valid, keys, rows=None):
    import ctypes as C
    from cases import value_type_c
    hb = OracleBackend().host(Batch(types, cols, valid))
    tt = (O.abi.QdbValueType * len(types))(*[value_type_c(t) for t in types])
    kk = (C.c_int32 * len(keys))(*keys)
    n = len(cols[0]) if rows is None else rows
    return [int(O.LIB.oracle_key_hash(hb.rs.Columns, tt, kk, len(keys), i)) for i in range(n)]


def test_key_hash_known_answers():
    """Values derived from the generator's text, not from the oracle:
    - BIGINT NOT NULL k = 1: struct {key_0}; h = 0 ^ (mix(1) + 0x9E3779B97F4A7C15)
      = 0xe61c480508b75932;
    - nullable BIGINT 5: struct {valid_0, key_0}; valid_0 = true hashes as
      1 * 2685821657736338717, then key_0 = mix(5): 0x33d914a28e354a67;
    - nullable BIGINT NULL: both field hashes 0: 0xcd94bf3ecef65c4f."""
    assert _key_hash([T(SIGNED, 8)], [np.array([1], np.int64)], None, [0]) == [0xE61C480508B75932]
    got = _key_hash([T(SIGNED, 8, nullable=True)], [np.array([5, 5], np.int64)], [np.array([True, False])], [0])
    assert got == [0x33D914A28E354A67, 0xCD94BF3ECEF65C4F]


def _py_mix(h):
    M = (1 << 64) - 1
    h ^= h >> 12
    h ^= (h << 25) & M
    h ^= h >> 27
    return (h * 2685821657736338717) & M


def _py_combine(h, f):
    M = (1 << 64) - 1
    return (h ^ ((f + 0x9E3779B97F4A7C15 + ((h << 6) & M) + (h >> 2)) & M)) & M


def test_key_hash_struct_rules():
    """Each rule of the generator checked on its own: the zero-extension of a
    narrow signed key (i8 -1 hashes the bits 0xff), the i128 leaf's lo/hi
    combine (gen.cpp:281-306), the bool leaf without a mix, and the ordered
    combine of two fields."""
    assert _key_hash([T(SIGNED, 1)], [np.array([-1], np.int8)], None, [0]) == [_py_combine(0, _py_mix(0xFF))]
    v = -(1 << 100) + 12345
    u = v & ((1 << 128) - 1)
    lo, hi = _py_mix(u & ((1 << 64) - 1)), _py_mix(u >> 64)
    leaf = lo ^ ((hi + 0x9E3779B97F4A7C15 + ((lo << 6) & ((1 << 64) - 1)) + (lo >> 2)) & ((1 << 64) - 1))
    assert _key_hash([T(DECIMAL, 16, 2)], [[v]], None, [0]) == [_py_combine(0, leaf)]
    assert _key_hash([T(3, 1)], [np.array([True, False])], None, [0]) == [
        _py_combine(0, 2685821657736338717), _py_combine(0, 0)]
    two = _key_hash([T(SIGNED, 4), T(SIGNED, 8)], [np.array([7], np.int32), np.array([9], np.int64)], None, [0, 1])
    assert two == [_py_combine(_py_combine(0, _py_mix(7)), _py_mix(9))]
    # Field order matters (the combine is ordered).
    rev = _key_hash([T(SIGNED, 4), T(SIGNED, 8)], [np.array([7], np.int32), np.array([9], np.int64)], None, [1, 0])
    assert rev != two


def test_key_hash_invariants():
    """NULL differs from every value of a nullable key (the valid_ field
    enters the combine); +0 and -0 hash alike, as does every NaN bit pattern
    (canonical float bits, gen.cpp:192-231); a non-null value of a nullable
    key differs from the same value of a NOT NULL key."""
    vals = np.array([0, 1, -1, 5], np.int64)
    nul = _key_hash([T(SIGNED, 8, nullable=True)], [vals], [np.array([False] * 4)], [0])
    val = _key_hash([T(SIGNED, 8, nullable=True)], [vals], [np.array([True] * 4)], [0])
    assert len(set(nul)) == 1 and not set(nul) & set(val)
    assert val != _key_hash([T(SIGNED, 8)], [vals], None, [0])
    nans = np.array([0.0, -0.0, np.nan, -np.nan, np.frombuffer(np.uint64(0x7FF0000000000123).tobytes(), np.float64)[0], 1.5])
    h = _key_hash([T(FLOAT, 8)], [nans], None, [0])
    assert h[0] == h[1] and h[2] == h[3] == h[4] and len({h[0], h[2], h[5]}) == 3


# ---------------------------------------------------------------------------
# Runtime join filter (qdb/kernel/join/runtime_filter.oz)
# ---------------------------------------------------------------------------

def test_rf_mix_is_splitmix64():
    """rf_mix is the splitmix64 finalizer: its first three outputs from seed
    0 (the state stepped by 0x9E3779B97F4A7C15) are the published
    0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f."""
    g = 0x9E3779B97F4A7C15
    M = (1 << 64) - 1
    assert [O.LIB.oracle_rf_mix((k * g) & M) for k in (1, 2, 3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                                      0x06C45D188009454F]


def test_rf_exact_set_is_membership():
    """Up to MaxKeys distinct hashes the filter is exact: a probe keeps a
    selected row iff its hash is one of the selected build hashes (hash 0
    aside, which the table stores as 1)."""
    rng = np.random.default_rng(2)
    build = rng.integers(1, 1 << 63, 5000, dtype=np.uint64)
    bsel = (rng.random(5000) > 0.3).astype(np.uint8)
    f = O.OracleRuntimeFilter(max_keys=1 << 14)
    f.build(np.concatenate([build, build[:100]]), np.concatenate([bsel, bsel[:100]]))
    assert f.publish() == 1
    probe = np.concatenate([build, rng.integers(1, 1 << 63, 20000, dtype=np.uint64)])
    psel = (rng.random(len(probe)) > 0.1).astype(np.uint8)
    out, kept = f.probe(probe, psel)
    want = np.isin(probe, build[bsel.astype(bool)]) & psel.astype(bool)
    assert np.array_equal(out.astype(bool), want) and kept == int(want.sum())


def test_rf_exact_counts_raw_hashes():
    """The exact set's keys are raw hashes: 0 and 1 are two distinct keys
    (both stored as 1 in the table, at their own slots), so a probe of either
    is kept, and they count as two against MaxKeys."""
    f = O.OracleRuntimeFilter(max_keys=1)
    f.build(np.array([0, 1], np.uint64))
    assert f.publish() == 2
    f = O.OracleRuntimeFilter(max_keys=2)
    f.build(np.array([0, 1], np.uint64))
    assert f.publish() == 1
    out, kept = f.probe(np.array([0, 1, 2], np.uint64))
    assert list(out) == [1, 1, 0] and kept == 2


def test_rf_switches_to_bloom_past_max_keys():
    """MaxKeys distinct hashes keep the exact set; one more makes it a Bloom
    filter, which never rejects a build hash and rejects most others; with
    under 64 bytes of Bloom storage the filter cannot answer and keeps every
    selected row."""
    rng = np.random.default_rng(3)
    keys = rng.integers(1, 1 << 63, 1001, dtype=np.uint64)
    f = O.OracleRuntimeFilter(max_keys=1000)
    f.build(keys[:1000])
    assert f.publish() == 1
    f = O.OracleRuntimeFilter(max_keys=1000, bloom_bytes=1 << 16)
    f.build(keys)
    assert f.publish() == 2
    out, kept = f.probe(keys)
    assert out.all() and kept == len(keys)
    others = rng.integers(1, 1 << 63, 100000, dtype=np.uint64)
    _, fp = f.probe(others)
    # 1001 hashes in 1024 blocks of 512 bits, 8 bits each: a false positive
    # needs 8 set bits in one block, about (1 - e^-(8 / 512))^8 ~ 3e-15 per
    # lightly filled block; a handful at most.
    assert fp < 100
    f = O.OracleRuntimeFilter(max_keys=10, bloom_bytes=32)
    f.build(keys)
    assert f.publish() == 0
    sel = (rng.random(len(others)) > 0.5).astype(np.uint8)
    out, kept = f.probe(others, sel)
    assert np.array_equal(out, sel) and kept == int(sel.sum())


# ---------------------------------------------------------------------------
# Hash shuffle and partial aggregate
# ---------------------------------------------------------------------------

def _rows_per_output(rows, offsets):
    return [rows[offsets[o]:offsets[o + 1]].tolist() for o in range(len(offsets) - 1)]


def test_shuffle_reference_goldens():
    """The reference's routing cases (test/test_scheduler_runtime_adapter.cpp:
    1496-1536, 1538-1562): hashes 0, 1, 2, 3, 5 over two outputs, without and
    with the routing table 1, 0, 0, 1; a partition routed nowhere fails."""
    h = np.array([0, 1, 2, 3, 5], np.uint64)
    assert _rows_per_output(*O.shuffle_partition(h, None, 2)) == [[0, 2], [1, 3, 4]]
    assert _rows_per_output(*O.shuffle_partition(h, None, 2, routes=[1, 0, 0, 1])) == [[1, 2, 4], [0, 3]]
    with pytest.raises(ValueError, match="partition 1, which it routes nowhere"):
        O.shuffle_partition(np.array([0, 3], np.uint64), None, 2, routes=[0, 0xFFFFFFFF])


def test_shuffle_partition_properties():
    """Each selected row lands in exactly one output, in input order, at its
    hash's residue (checked with Python integers); 8 outputs take the low 3
    bits; an identity routing table is no table, a permuted one permutes the
    outputs, and a table of 12 partitions folded onto 3 outputs equals 3
    plain outputs."""
    rng = np.random.default_rng(5)
    n = 3000
    h = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    h[:4] = [0, (1 << 64) - 1, 1 << 63, 12345]
    sel = rng.random(n) > 0.3
    for outputs in (1, 3, 8, 257):
        rows, offs = O.shuffle_partition(h, sel, outputs)
        assert sorted(rows.tolist()) == np.flatnonzero(sel).tolist()
        for o, part in enumerate(_rows_per_output(rows, offs)):
            assert part == sorted(part)
            assert all(int(h[r]) - (int(h[r]) // outputs) * outputs == o for r in part)
    for o, part in enumerate(_rows_per_output(*O.shuffle_partition(h, None, 8))):
        assert all(int(h[r]) & 7 == o for r in part)
    plain = _rows_per_output(*O.shuffle_partition(h, sel, 5))
    assert _rows_per_output(*O.shuffle_partition(h, sel, 5, routes=list(range(5)))) == plain
    perm = [3, 0, 4, 1, 2]
    routed = _rows_per_output(*O.shuffle_partition(h, sel, 5, routes=perm))
    assert all(routed[perm[p]] == plain[p] for p in range(5))
    assert _rows_per_output(*O.shuffle_partition(h, sel, 3, routes=[i % 3 for i in range(12)])) == \
        _rows_per_output(*O.shuffle_partition(h, sel, 3))


def _partial_types(spec_c):
    return [T(t.Kind, t.Width, t.Scale, bool(t.Nullable), t.Precision) for t in O.output_types(spec_c)]


def _combine_case(case: AggCase, ptypes):
    nk = len(case.keys)
    aggs = [(AGG_SUM if f in (AGG_COUNT, AGG_COUNT_STAR) else f, nk + i) for i, (f, _) in enumerate(case.aggs)]
    return AggCase(ptypes, list(range(nk)), aggs)


def _as_batch(ptypes, columns):
    cols, valid = [], []
    for t, (vals, v) in zip(ptypes, columns):
        if t.width == 16:
            vals = [int(x) for x in vals]
        elif t.kind == FLOAT:
            vals = np.asarray(vals, np.float64)
        else:
            vals = np.asarray(vals).astype(np.int64 if t.width == 8 else vals.dtype)
        cols.append(vals)
        valid.append(None if v is None else np.asarray(v, bool))
    return Batch(ptypes, cols, valid if any(x is not None for x in valid) else None)


def _partial_cases():
    types = [T(SIGNED, 8), T(SIGNED, 4, nullable=True), T(DECIMAL, 16, 2, True, 30), T(FLOAT, 8)]
    return types, AggCase(types, [0], [(AGG_COUNT_STAR, -1), (AGG_COUNT, 1), (AGG_SUM, 1), (AGG_SUM, 2), (AGG_MIN, 3),
                                       (AGG_MAX, 1)])


def _partial_batches(rng, types, nbatches, rows, distinct):
    out = []
    for _ in range(nbatches):
        cols = [rng.integers(0, distinct, rows).astype(np.int64), rng.integers(-50, 50, rows).astype(np.int32),
                [int(x) * 10 ** 20 + 3 for x in rng.integers(-10 ** 6, 10 ** 6, rows)], rng.normal(0, 1, rows)]
        valid = [None, rng.random(rows) > 0.2, rng.random(rows) > 0.3, None]
        out.append(Batch(types, cols, valid, (rng.random(rows) > 0.1).astype(np.uint8)))
    return out


@pytest.mark.parametrize("distinct,sample,max_groups,want_passing,want_tables", [
    (40, 5000, 1 << 20, False, 0),        # reduces well: one table, kept to the end
    (10 ** 9, 5000, 1 << 20, True, 1),    # unique keys: the table is emitted at 6000 rows, then states
    (2500, 10 ** 9, 1500, False, None),   # bounded table: emits whenever it reaches 1500 groups
])
def test_partial_aggregate_policy_and_cascade(distinct, sample, max_groups, want_passing, want_tables):
    """The partial aggregate's decisions (TPartialAggregateProcessor::Add),
    and the cascade's invariant that makes them safe: the final aggregate
    (COUNT -> SUM) over whatever the partial phase emitted, tables and
    single-row states alike, equals the one-phase aggregate."""
    rng = np.random.default_rng(distinct % 1000 + 7)
    types, case = _partial_cases()
    batches = _partial_batches(rng, types, 6, 2000, distinct)
    be = OracleBackend()
    spec = case.to_c()
    pa = O.PartialAggregate(spec, sample_rows=sample, max_groups=max_groups)
    for b in batches:
        pa.add(be.host(b))
    assert pa.passing == want_passing
    tables = [o for kind, o in pa.outputs if kind == "table"]
    if want_tables is not None:
        assert len(tables) == want_tables
    else:
        assert len(tables) >= 2 and all(len(t[0][0]) >= max_groups for t in tables)
    ptypes = _partial_types(spec)
    parts = [_as_batch(ptypes, t) for t in tables]
    for kind, hb in pa.outputs:
        if kind == "states":
            parts.append(_as_batch(ptypes, O.partial_states(spec, hb)))
    if not pa.passing and pa.table:
        parts.append(_as_batch(ptypes, O.aggregate(spec, pa.table)))
    final = sorted(result_rows(be.aggregate(_combine_case(case, ptypes), parts)), key=sort_key)
    direct = sorted(result_rows(be.aggregate(case, batches)), key=sort_key)
    assert final == direct


CMP_PY = [("EQ", lambda a, b: a == b), ("NE", lambda a, b: a != b), ("LT", lambda a, b: a < b),
          ("LE", lambda a, b: a <= b), ("GT", lambda a, b: a > b), ("GE", lambda a, b: a >= b)]


@pytest.mark.parametrize("name,fn", CMP_PY, ids=[n for n, _ in CMP_PY])
def test_comparisons_against_python(name, fn):
    """Every comparison op of the filter language, pinned against Python's
    own comparison of the same values under SQL's rule that a comparison
    with NULL is NULL (and NULL is not TRUE): int64 columns with NULLs and
    equal pairs, DECIMAL(15,2) against DECIMAL(15,4) after the scale
    alignment (compared as exact fractions), and f64 without NaN (with -0
    equal to +0); then NOT and AND/OR over the comparison's NULLs."""
    from fractions import Fraction
    E = cases._abi()
    op = getattr(E, "CMP_" + name)
    rng = np.random.default_rng(hash(name) % 1000)
    n = 4000
    x = rng.integers(-5, 6, n).astype(np.int64)
    y = rng.integers(-5, 6, n).astype(np.int64)
    xv, yv = rng.random(n) > 0.2, rng.random(n) > 0.2
    d1 = rng.integers(-500, 500, n).astype(np.int64)             # DECIMAL(15,2)
    d2 = (d1 * 100 + rng.integers(-1, 2, n) * rng.integers(0, 2, n)).astype(np.int64)  # DECIMAL(15,4), often equal
    f = rng.choice([-1.5, -0.0, 0.0, 2.25, 1e300, -1e-300], n)
    g = rng.choice([-1.5, 0.0, -0.0, 2.25, 1e300, -1e-300], n)
    types = [T(SIGNED, 8, nullable=True), T(SIGNED, 8, nullable=True), T(DECIMAL, 8, 2, False, 15), T(DECIMAL, 8, 4, False, 15),
             T(FLOAT, 8), T(FLOAT, 8)]
    batch = Batch(types, [x, y, d1, d2, f, g], [xv, yv, None, None, None, None])
    ob = OracleBackend()

    def run(build):
        p = cases.Prog(types)
        r = build(p)
        return ob.filter(types, p, r, batch).astype(bool)

    # int64 with NULLs
    got = run(lambda p: p.cmp(op, p.col(0), p.col(1)))
    want = np.array([bool(xv[i] and yv[i] and fn(int(x[i]), int(y[i]))) for i in range(n)])
    assert np.array_equal(got, want)
    assert want.any() and (~want).any()
    # decimals at different scales
    got = run(lambda p: p.cmp(op, p.scale_up(p.col(2), 2, T(DECIMAL, 8, 4, False, 17)), p.col(3)))
    want = np.array([fn(Fraction(int(d1[i]), 100), Fraction(int(d2[i]), 10000)) for i in range(n)])
    assert np.array_equal(got, want)
    # f64 (no NaN)
    got = run(lambda p: p.cmp(op, p.col(4), p.col(5)))
    assert np.array_equal(got, np.array([fn(float(f[i]), float(g[i])) for i in range(n)]))
    # NOT of a comparison with NULLs: NULL stays NULL (not selected)
    got = run(lambda p: p._emit(E.EXPR_NOT, E.REG_BOOL, T(3, 1, nullable=True), a=p.cmp(op, p.col(0), p.col(1))))
    want = np.array([bool(xv[i] and yv[i] and not fn(int(x[i]), int(y[i]))) for i in range(n)])
    assert np.array_equal(got, want)
    # AND / OR with a NULL operand: FALSE AND NULL = FALSE, TRUE OR NULL = TRUE
    def and3(a, b):
        return False if (a is False or b is False) else (None if (a is None or b is None) else True)

    def or3(a, b):
        return True if (a is True or b is True) else (None if (a is None or b is None) else False)

    for lop, py in ((E.EXPR_AND, and3), (E.EXPR_OR, or3)):
        got = run(lambda p: p._emit(lop, E.REG_BOOL, T(3, 1, nullable=True), a=p.cmp(op, p.col(0), p.col(1)),
                                    b=p.cmp(E.CMP_GE, p.col(4), p.const_f64(0.0))))
        want = []
        for i in range(n):
            a = fn(int(x[i]), int(y[i])) if (xv[i] and yv[i]) else None
            b = float(f[i]) >= 0.0
            r = py(a, b)
            want.append(r is True)
        assert np.array_equal(got, np.array(want))
